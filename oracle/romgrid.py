"""The ROM runs' data-grid map restated in Python (test infrastructure only).

GRID_FUNCTIONS (MainSrc/GRID_FUNCTIONS.f08):
* GRIDMAP_GEN_AVG / _EDGV / _EDGH / _LR_BND / _BT_BND (:49-200): point coordinates (x, y)
  of cell centres, V faces, H faces and boundary faces, x fastest.
* GENERATE_GRIDS_XY (:689-752): the simulation grids and the data grids with a zero-width
  cell appended on every side (so data points sit on the domain boundary), and MAP_GRIDS
  between them.
* MAP_GRIDS (:222-327): for every simulation point the closest data point (first in
  order on ties) and, from where it lies relative to the point, the four data points of
  the enclosing cell: l1 bottom left, l2 bottom right, l3 top left, l4 top right, as
  positions in the coordinate array (odd, 1-based) and as point indices (VMap).
* BILINEAR_INTERPOLATE (:366-400) and FMAP (:14-43): the bilinear interpolant per group.
* GENERATE_GRIDS_T (:754-817): the data time of each simulation time (TMap).
* COLLECT_LR_BND / _BT_BND / _CRN_BND, EXTND_BT_BND / _LR_BND, AVG_APND_BND (:406-659): the
  boundary values appended to the data fields before the map, as POD_RECONSTRUCT_fg uses
  them (MainSrc/POD_ROUTINES.f08:28-236); the side extension of the C_g boundary vector
  (POD_RECONSTRUCT_BCg, :253-353).

Parity pinning (tests/test_oracle_romgrid.py): FMAP through MAP_GRIDS equals scipy's
RegularGridInterpolator (bilinear on the rectilinear data grid, an independent method) at
every simulation point; identical grids give the data back exactly; a field bilinear in x
and y is reproduced exactly; TMap brackets every simulation time by its data times.
"""
from __future__ import annotations

import math

import numpy as np


# ---------------------------------------------------------------- point grids (x fastest)

def gridmap_avg(delx, dely):
    nx, ny = len(delx), len(dely)
    xloc = np.empty(nx)
    xloc[0] = delx[0] / 2.0
    for i in range(1, nx):
        xloc[i] = xloc[i - 1] + (delx[i - 1] + delx[i]) / 2.0
    out, y = [], 0.0
    for j in range(ny):
        y = y + dely[j] / 2.0
        for i in range(nx):
            out += [xloc[i], y]
        y = y + dely[j] / 2.0
    return np.array(out)


def gridmap_edgv(delx, dely, nxp):
    """nxp = number of x points (N_x + 1)."""
    xloc = np.empty(nxp)
    xloc[0] = 0.0
    for i in range(1, nxp):
        xloc[i] = xloc[i - 1] + delx[i - 1]
    out, y = [], 0.0
    for j in range(len(dely)):
        y = y + dely[j] / 2.0
        for i in range(nxp):
            out += [xloc[i], y]
        y = y + dely[j] / 2.0
    return np.array(out)


def gridmap_edgh(delx, dely, nyp):
    """nyp = number of y points (N_y + 1)."""
    nx = len(delx)
    xloc = np.empty(nx)
    xloc[0] = delx[0] / 2.0
    for i in range(1, nx):
        xloc[i] = xloc[i - 1] + (delx[i - 1] + delx[i]) / 2.0
    out, y = [], 0.0
    for j in range(nyp):
        for i in range(nx):
            out += [xloc[i], y]
        if j < len(dely):
            y = y + dely[j]
    return np.array(out)


def gridmap_lr(x, dely):
    out, y = [], 0.0
    for d in dely:
        y = y + d / 2.0
        out += [x, y]
        y = y + d / 2.0
    return np.array(out)


def gridmap_bt(delx, y):
    out, x = [], 0.0
    for d in delx:
        x = x + d / 2.0
        out += [x, y]
        x = x + d / 2.0
    return np.array(out)


# ---------------------------------------------------------------- MAP_GRIDS

def map_grids(grid1, grid2, xlen1, len1=None, len2=None):
    """Map (1-based positions in grid1's coordinate array, 4 per grid2 point) and VMap.
    len1 / len2: coordinates of grid1 searched / of grid2 mapped (the whole grids unless
    given; the neighbours l2..l4 may lie beyond len1)."""
    len1 = len(grid1) if len1 is None else len1
    len2 = len(grid2) if len2 is None else len2
    out = []
    tol = 1e-10
    for j in range(0, len2, 2):
        x2, y2 = grid2[j], grid2[j + 1]
        d1, l1 = math.inf, None
        for i in range(0, len1, 2):
            r = abs(math.sqrt((grid1[i] - x2) ** 2 + (grid1[i + 1] - y2) ** 2))
            if r < d1:
                d1, l1 = r, i + 1          # 1-based position of the x coordinate
        gx, gy = grid1[l1 - 1], grid1[l1]
        if gx - x2 < -tol:
            if gy - y2 < -tol:                       # bottom left
                l = [l1, l1 + 2, l1 + xlen1, l1 + xlen1 + 2]
            elif gy - y2 > tol:                      # top left
                l3 = l1
                a = l3 - xlen1
                l = [a, a + 2, l3, l3 + 2]
            else:                                    # exact left
                l = [l1, l1 + 2, l1, l1 + 2]
        elif gx - x2 > tol:
            if gy - y2 < -tol:                       # bottom right
                l2 = l1
                a = l2 - 2
                l = [a, l2, a + xlen1, a + xlen1 + 2]
            elif gy - y2 > tol:                      # top right
                l4 = l1
                l2 = l4 - xlen1
                l = [l2 - 2, l2, l4 - 2, l4]
            else:                                    # exact right
                l2 = l1
                a = l2 - 2
                l = [a, l2, a, l2]
        else:
            if gy - y2 < -tol:                       # exact bottom
                l = [l1, l1, l1 + xlen1, l1 + xlen1]
            elif gy - y2 > tol:                      # exact top
                a = l1 - xlen1
                l = [a, a, l1, l1]
            else:                                    # exact point
                l = [l1, l1, l1, l1]
        out += l
    gmap = np.array(out, dtype=np.int64)
    return gmap, (gmap + 1) // 2


def bilinear(f11, f12, f21, f22, x, x1, x2, y, y1, y2):
    if (x2 - x1) == 0:
        if (y2 - y1) == 0:
            return f11
        yy1, yy2 = y - y1, y2 - y
        return (f11 * yy2 + f12 * yy1) / (y2 - y1)
    if (y2 - y1) == 0:
        xx1, xx2 = x - x1, x2 - x
        return (f11 * xx2 + f21 * xx1) / (x2 - x1)
    xx1, xx2, yy1, yy2 = x - x1, x2 - x, y - y1, y2 - y
    return (f11 * xx2 * yy2 + f21 * xx1 * yy2 + f12 * xx2 * yy1 + f22 * xx1 * yy1) / ((x2 - x1) * (y2 - y1))


def fmap(f1, grid1, grid2, gmap, vmap, glen1, glen2, ngroups):
    """FMAP: f1 holds ngroups blocks of glen1 values on grid1; returns ngroups blocks of glen2."""
    f1 = np.asarray(f1, dtype=np.float64).ravel(order="F")
    out = np.empty(ngroups * glen2)
    for g in range(ngroups):
        p1 = g * glen1
        for j in range(glen2):
            v = vmap[4 * j:4 * j + 4] - 1
            gm = gmap[4 * j:4 * j + 4] - 1
            out[g * glen2 + j] = bilinear(f1[p1 + v[0]], f1[p1 + v[2]], f1[p1 + v[1]], f1[p1 + v[3]],
                                          grid2[2 * j], grid1[gm[0]], grid1[gm[3]],
                                          grid2[2 * j + 1], grid1[gm[0] + 1], grid1[gm[3] + 1])
    return out


# ---------------------------------------------------------------- GENERATE_GRIDS

class Grids:
    """GENERATE_GRIDS_XY's grids and maps for simulation cells (delx, dely) and data
    cells (delx_d, dely_d); xlen_d, ylen_d, xlen, ylen as the reference passes them."""

    def __init__(self, delx_d, dely_d, delx, dely, xlen_d, ylen_d, xlen, ylen):
        delx_d, dely_d = np.asarray(delx_d, float), np.asarray(dely_d, float)
        delx, dely = np.asarray(delx, float), np.asarray(dely, float)
        dnx, dny, nx, ny = len(delx_d), len(dely_d), len(delx), len(dely)
        self.dnx, self.dny, self.nx, self.ny = dnx, dny, nx, ny
        dx2 = np.concatenate([[0.0], delx_d, [0.0]])
        dy2 = np.concatenate([[0.0], dely_d, [0.0]])
        self.sim_avg = gridmap_avg(delx, dely)
        self.sim_edgv = gridmap_edgv(delx, dely, nx + 1)
        self.sim_edgh = gridmap_edgh(delx, dely, ny + 1)
        # the right side at x = ylen and the top at y = xlen, as GENERATE_GRIDS_XY passes them
        self.sim_bnds = [gridmap_lr(0.0, dely), gridmap_bt(delx, 0.0), gridmap_lr(ylen, dely), gridmap_bt(delx, xlen)]
        self.dat_avg = gridmap_avg(dx2, dy2)
        self.dat_edgv = gridmap_edgv(delx_d, dy2, dnx + 1)
        self.dat_edgh = gridmap_edgh(dx2, dely_d, dny + 1)
        self.dat_bnds = [gridmap_lr(0.0, dy2), gridmap_bt(dx2, 0.0), gridmap_lr(ylen_d, dy2), gridmap_bt(dx2, xlen_d)]
        self.map_avg = map_grids(self.dat_avg, self.sim_avg, 2 * (dnx + 2))
        self.map_edgv = map_grids(self.dat_edgv, self.sim_edgv, 2 * (dnx + 1))
        self.map_edgh = map_grids(self.dat_edgh, self.sim_edgh, 2 * (dnx + 2))
        # the four sides as GENERATE_GRIDS_XY maps them (:741-750): 2 dN_y + 2 data coordinates
        # searched and 2 N_y simulation coordinates mapped on every side (the last data point,
        # on the far corner, is reached only as a neighbour); bottom and top need N_x = N_y
        # and dN_x = dN_y for that to cover them
        square = dnx == dny and nx == ny
        self.map_bnds = [map_grids(self.dat_bnds[s], self.sim_bnds[s], 2 if s in (0, 2) else 0, 2 * dny + 2, 2 * ny)
                         for s in range(4)] if square else None


def generate_grids_t(delt_d, delt, start_time_d, start_time, dnt, nt):
    """TMap (1-based data-time pairs per simulation time), Sim_TGrid, Dat_TGrid."""
    dat = np.empty(dnt)
    dat[0] = start_time_d + delt_d
    for j in range(1, dnt):
        dat[j] = dat[j - 1] + delt_d
    sim = np.empty(nt)
    sim[0] = start_time + delt
    for j in range(1, nt):
        sim[j] = sim[j - 1] + delt
    tmap = []
    for i in range(nt):
        if dat[0] >= sim[i]:
            tmap += [1, 1]
        elif dat[dnt - 1] <= sim[i]:
            tmap += [dnt, dnt]
        else:
            for j in range(2, dnt + 1):
                if abs(dat[j - 1] - sim[i]) / dat[j - 1] <= 1e-14:
                    tmap += [j, j]
                    break
                if dat[j - 1] > sim[i]:
                    tmap += [j, j - 1]
                    break
    return np.array(tmap), sim, dat


# ---------------------------------------------------------------- boundary values

def _g(f, shape):
    return np.asarray(f, dtype=np.float64).reshape(shape, order="F")


def extnd_bt(f, nxp, ny, ng):
    """EXTND_BT_BND: rows 0 and N_y-1 repeated below and above (nxp points per row)."""
    a = _g(f, (nxp, ny, ng))
    return np.concatenate([a[:, :1], a, a[:, -1:]], axis=1).ravel(order="F")


def extnd_lr(f, nx, nyp, ng):
    """EXTND_LR_BND: columns 0 and N_x-1 repeated left and right."""
    a = _g(f, (nx, nyp, ng))
    return np.concatenate([a[:1], a, a[-1:]], axis=0).ravel(order="F")


def avg_apnd(f, lr, bt, crn, nx, ny, ng):
    """AVG_APND_BND: lr (2, N_y, N_g) left/right values per row, bt (N_x, 2, N_g) bottom/top
    rows, crn (4, N_g) corners bottom-left, bottom-right, top-left, top-right."""
    a = _g(f, (nx, ny, ng))
    out = np.empty((nx + 2, ny + 2, ng))
    out[1:-1, 1:-1] = a
    out[0, 1:-1], out[-1, 1:-1] = lr[0], lr[1]
    out[1:-1, 0], out[1:-1, -1] = bt[:, 0], bt[:, 1]
    out[0, 0], out[-1, 0], out[0, -1], out[-1, -1] = crn[0], crn[1], crn[2], crn[3]
    return out.ravel(order="F")


def collect_lr(f, nxp, ny, ng):
    a = _g(f, (nxp, ny, ng))
    return np.stack([a[0], a[-1]])                    # (2, N_y, N_g)


def collect_bt(f, nx, nyp, ng):
    a = _g(f, (nx, nyp, ng))
    return np.stack([a[:, 0], a[:, -1]], axis=1)      # (N_x, 2, N_g)


def collect_crn(f, nx, ny, ng):
    a = _g(f, (nx, ny, ng))
    return np.stack([a[0, 0], a[-1, 0], a[0, -1], a[-1, -1]])


def map_fields(G: Grids, dat, ng):
    """POD_RECONSTRUCT_fg after the reconstructions and POD_RECONSTRUCT_BCg after its own:
    dat holds the data-grid fields fg_avg_xx, fg_avg_yy, fg_edgV_xx, fg_edgV_xy, fg_edgH_yy,
    fg_edgH_xy and BCg; returns the simulation-grid fields and Cg_L/B/R/T."""
    dnx, dny, nx, ny = G.dnx, G.dny, G.nx, G.ny
    out = {}
    xxbnd = collect_lr(dat["fg_edgV_xx"], dnx + 1, dny, ng)
    for n in ("fg_edgV_xx", "fg_edgV_xy"):
        f2 = extnd_bt(dat[n], dnx + 1, dny, ng)
        out[n] = fmap(f2, G.dat_edgv, G.sim_edgv, *G.map_edgv, (dnx + 1) * (dny + 2), (nx + 1) * ny, ng)
    yybnd = collect_bt(dat["fg_edgH_yy"], dnx, dny + 1, ng)
    for n in ("fg_edgH_yy", "fg_edgH_xy"):
        f2 = extnd_lr(dat[n], dnx, dny + 1, ng)
        out[n] = fmap(f2, G.dat_edgh, G.sim_edgh, *G.map_edgh, (dnx + 2) * (dny + 1), nx * (ny + 1), ng)
    a = dat["fg_avg_xx"]
    f2 = avg_apnd(a, xxbnd, collect_bt(a, dnx, dny, ng), collect_crn(a, dnx, dny, ng), dnx, dny, ng)
    out["fg_avg_xx"] = fmap(f2, G.dat_avg, G.sim_avg, *G.map_avg, (dnx + 2) * (dny + 2), nx * ny, ng)
    a = dat["fg_avg_yy"]
    f2 = avg_apnd(a, collect_lr(a, dnx, dny, ng), yybnd, collect_crn(a, dnx, dny, ng), dnx, dny, ng)
    out["fg_avg_yy"] = fmap(f2, G.dat_avg, G.sim_avg, *G.map_avg, (dnx + 2) * (dny + 2), nx * ny, ng)
    # C_g: per group left (dN_y), bottom (dN_x), right (dN_y), top (dN_x), each side extended by
    # its end values, then mapped onto the simulation sides
    bc = _g(dat["BCg"], (2 * (dnx + dny), ng))
    sides = [(0, dny), (dny, dnx), (dny + dnx, dny), (2 * dny + dnx, dnx)]
    names = ("Cg_L", "Cg_B", "Cg_R", "Cg_T")
    for s, ((o, n), name) in enumerate(zip(sides, names)):
        v = bc[o:o + n]
        ext = np.concatenate([v[:1], v, v[-1:]], axis=0).ravel(order="F")
        glen2 = ny if s in (0, 2) else nx
        out[name] = fmap(ext, G.dat_bnds[s], G.sim_bnds[s], *G.map_bnds[s], n + 2, glen2, ng)
    return out
