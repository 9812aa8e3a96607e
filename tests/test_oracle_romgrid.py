"""Pins of the data-grid map restatement (oracle/romgrid.py; MainSrc/GRID_FUNCTIONS.f08 and the
boundary appends of POD_RECONSTRUCT_fg, MainSrc/POD_ROUTINES.f08:28-353), CPU only.

* FMAP through MAP_GRIDS is bilinear interpolation on the (rectilinear) data grid: it equals
  scipy's RegularGridInterpolator, an independent implementation, at every simulation point,
  on non-uniform grids, refining and coarsening, for every layout.
* Identical data and simulation grids give every field back exactly.
* The boundary values the reference appends make the map exact for fields linear across the
  boundary they come from: fg_avg_xx takes its left/right values from fg_edgV_xx's boundary
  faces, fg_avg_yy its bottom/top values from fg_edgH_yy's, the face fields repeat their last
  row or column.  A field F = a + b x sampled where each array lives is mapped exactly at every
  simulation point away from the corners (F = a + c y for the y-sides).
* TMap brackets every simulation time between its two data times (or clamps to the ends).
"""
import numpy as np
import pytest
from scipy.interpolate import RegularGridInterpolator

from oracle import romgrid as R


def _grids(dnx, dny, nx, ny, seed):
    rng = np.random.default_rng(seed)
    dxd = rng.uniform(0.5, 1.5, dnx)
    dyd = rng.uniform(0.5, 1.5, dny)
    dxd *= 1.0 / dxd.sum()
    dyd *= 1.0 / dyd.sum()
    dx = np.full(nx, 1.0 / nx)
    dy = np.full(ny, 1.0 / ny)
    return dxd, dyd, dx, dy


def _axes(grid, npx):
    pts = grid.reshape(-1, 2)
    return pts[:npx, 0], pts[::npx, 1]


@pytest.mark.parametrize("dn,n", [((7, 6), (11, 9)), ((12, 10), (5, 4)), ((8, 8), (8, 8))])
def test_fmap_is_bilinear_interpolation(dn, n):
    dnx, dny = dn
    nx, ny = n
    dxd, dyd, dx, dy = _grids(dnx, dny, nx, ny, dnx * 31 + nx)
    G = R.Grids(dxd, dyd, dx, dy, 1.0, 1.0, 1.0, 1.0)
    rng = np.random.default_rng(3)
    ng = 2
    for dat, sim, (gm, vm), npx, npy in ((G.dat_avg, G.sim_avg, G.map_avg, dnx + 2, dny + 2),
                                         (G.dat_edgv, G.sim_edgv, G.map_edgv, dnx + 1, dny + 2),
                                         (G.dat_edgh, G.sim_edgh, G.map_edgh, dnx + 2, dny + 1)):
        xs, ys = _axes(dat, npx)
        assert np.all(np.diff(xs) > 0) and np.all(np.diff(ys) > 0)
        f = rng.standard_normal((npx, npy, ng))
        got = R.fmap(f.ravel(order="F"), dat, sim, gm, vm, npx * npy, len(sim) // 2, ng)
        pts = sim.reshape(-1, 2)
        for g in range(ng):
            ref = RegularGridInterpolator((xs, ys), f[:, :, g], bounds_error=False, fill_value=None)(pts)
            np.testing.assert_allclose(got[g * len(pts):(g + 1) * len(pts)], ref, rtol=0, atol=1e-13)


def test_identical_grids_give_the_data_back():
    dxd, dyd, _, _ = _grids(9, 9, 9, 9, 5)
    G = R.Grids(dxd, dyd, dxd, dyd, 1.0, 1.0, 1.0, 1.0)
    rng = np.random.default_rng(8)
    ng, n = 3, 9
    shapes = {"fg_avg_xx": (n, n), "fg_avg_yy": (n, n), "fg_edgV_xx": (n + 1, n), "fg_edgV_xy": (n + 1, n),
              "fg_edgH_yy": (n, n + 1), "fg_edgH_xy": (n, n + 1), "BCg": (4 * n,)}
    dat = {k: rng.standard_normal(s + (ng,)).ravel(order="F") for k, s in shapes.items()}
    out = R.map_fields(G, dat, ng)
    for k in shapes:
        if k != "BCg":
            np.testing.assert_array_equal(out[k], dat[k], err_msg=k)
    bc = dat["BCg"].reshape(4 * n, ng, order="F")
    for s, name in enumerate(("Cg_L", "Cg_B", "Cg_R", "Cg_T")):
        np.testing.assert_array_equal(out[name], bc[s * n:(s + 1) * n].ravel(order="F"), err_msg=name)


def _sample(G, F, dxd, dyd, ng):
    """F(x, y) at every data array's points (cells, V faces, H faces, boundary faces)."""
    xc = np.cumsum(dxd) - dxd / 2
    yc = np.cumsum(dyd) - dyd / 2
    xe = np.concatenate([[0.0], np.cumsum(dxd)])
    ye = np.concatenate([[0.0], np.cumsum(dyd)])
    rep = lambda a: np.repeat(a[..., None], ng, axis=-1).ravel(order="F")
    d = {"fg_avg_xx": rep(F(*np.meshgrid(xc, yc, indexing="ij"))),
         "fg_avg_yy": rep(F(*np.meshgrid(xc, yc, indexing="ij"))),
         "fg_edgV_xx": rep(F(*np.meshgrid(xe, yc, indexing="ij"))),
         "fg_edgV_xy": rep(F(*np.meshgrid(xe, yc, indexing="ij"))),
         "fg_edgH_yy": rep(F(*np.meshgrid(xc, ye, indexing="ij"))),
         "fg_edgH_xy": rep(F(*np.meshgrid(xc, ye, indexing="ij")))}
    x1 = float(np.sum(dxd))
    d["BCg"] = rep(np.concatenate([F(0.0 * yc, yc), F(xc, 0.0 * xc), F(x1 + 0.0 * yc, yc), F(xc, x1 + 0.0 * xc)]))
    return d


@pytest.mark.parametrize("axis", ["x", "y"])
def test_boundary_appends_reproduce_linear_fields(axis):
    """A field that varies along one axis only is mapped exactly wherever the values the
    reference appends are samples of it: fg_avg_xx's left/right values are fg_edgV_xx's
    boundary faces (exact near x = 0 and x = xlen), fg_avg_yy's bottom/top values fg_edgH_yy's;
    rows and columns that are repeated (face fields, the other cell field's sides, C_g's side
    ends, the corners) are exact only across the axis the field does not vary along."""
    dnx = dny = 8
    nx = ny = 13
    dxd, dyd, dx, dy = _grids(dnx, dny, nx, ny, 17)
    G = R.Grids(dxd, dyd, dx, dy, 1.0, 1.0, 1.0, 1.0)
    F = (lambda x, y: 0.3 + 1.7 * x + 0.0 * y) if axis == "x" else (lambda x, y: -0.2 + 0.9 * y + 0.0 * x)
    ng = 2
    out = R.map_fields(G, _sample(G, F, dxd, dyd, ng), ng)
    xc = (np.arange(nx) + 0.5) / nx
    yc = (np.arange(ny) + 0.5) / ny
    xe = np.arange(nx + 1) / nx
    ye = np.arange(ny + 1) / ny
    # simulation points between the first and the last data cell centre
    inside = lambda v, d: (v > d[0] / 2) & (v < 1 - d[-1] / 2)
    along = (lambda xs, ys: np.outer(inside(xs, dxd), np.ones(len(ys), bool))) if axis == "x" else \
        (lambda xs, ys: np.outer(np.ones(len(xs), bool), inside(ys, dyd)))
    every = lambda xs, ys: np.ones((len(xs), len(ys)), bool)
    corners = lambda xs, ys: np.outer(inside(xs, dxd), np.ones(len(ys), bool)) | \
        np.outer(np.ones(len(xs), bool), inside(ys, dyd))
    own = "fg_avg_xx" if axis == "x" else "fg_avg_yy"       # takes the exact side values
    other = "fg_avg_yy" if axis == "x" else "fg_avg_xx"
    want = {own: (xc, yc, corners), other: (xc, yc, along),
            "fg_edgV_xx": (xe, yc, every if axis == "x" else along),
            "fg_edgV_xy": (xe, yc, every if axis == "x" else along),
            "fg_edgH_yy": (xc, ye, along if axis == "x" else every),
            "fg_edgH_xy": (xc, ye, along if axis == "x" else every)}
    for k, (xs, ys, mask) in want.items():
        got = out[k].reshape(len(xs), len(ys), ng, order="F")[:, :, 0]
        ref = F(*np.meshgrid(xs, ys, indexing="ij"))
        m = mask(xs, ys)
        assert m.sum() > 0.5 * m.size
        err = np.abs(got - ref)[m]
        assert err.max() < 1e-13, (k, axis, err.max())
    # C_g: sides along the axis repeat their end values
    side = {"Cg_L": (0 * yc, yc, axis == "x"), "Cg_B": (xc, 0 * xc, axis == "y"),
            "Cg_R": (1 + 0 * yc, yc, axis == "x"), "Cg_T": (xc, 1 + 0 * xc, axis == "y")}
    for k, (xs, ys, flat) in side.items():
        ref = F(xs, ys)
        got = out[k][:len(ref)]
        m = np.ones(len(ref), bool) if flat else inside(xs if k in ("Cg_B", "Cg_T") else ys,
                                                         dxd if k in ("Cg_B", "Cg_T") else dyd)
        assert np.abs(got - ref)[m].max() < 1e-13, k


def test_a_mistake_in_the_appends_is_caught():
    """The exactness above fails if fg_avg_xx took its side values from its own columns."""
    dnx = dny = 8
    dxd, dyd, dx, dy = _grids(dnx, dny, 13, 13, 17)
    G = R.Grids(dxd, dyd, dx, dy, 1.0, 1.0, 1.0, 1.0)
    F = lambda x, y: 0.3 + 1.7 * x + 0.0 * y
    d = _sample(G, F, dxd, dyd, 1)
    d_bad = dict(d)
    ev = d["fg_edgV_xx"].reshape(dnx + 1, dny, order="F").copy()
    av = d["fg_avg_xx"].reshape(dnx, dny, order="F")
    ev[0], ev[-1] = av[0], av[-1]          # side values = fg_avg_xx's own outer columns
    d_bad["fg_edgV_xx"] = ev.ravel(order="F")
    good = R.map_fields(G, d, 1)["fg_avg_xx"]
    bad = R.map_fields(G, d_bad, 1)["fg_avg_xx"]
    assert np.abs(good - bad).max() > 1e-3


def test_tmap_brackets_simulation_times():
    tm, sim, dat = R.generate_grids_t(0.3, 0.1, 0.0, 0.0, 5, 20)
    assert len(tm) == 40
    for i, s in enumerate(sim):
        a, b = tm[2 * i], tm[2 * i + 1]
        if s <= dat[0]:
            assert a == b == 1
        elif s >= dat[-1]:
            assert a == b == 5
        elif a == b:
            assert abs(dat[a - 1] - s) <= 1e-14 * dat[a - 1]
        else:
            assert b == a - 1 and dat[b - 1] < s < dat[a - 1]
