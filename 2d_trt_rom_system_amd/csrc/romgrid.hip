// The ROM runs' data grid: GENERATE_GRIDS_XYT (MainSrc/GRID_FUNCTIONS.f08:661-817) and the
// map of reconstructed fields from the data grid onto the simulation grid that
// POD_RECONSTRUCT_fg / DMD_RECONSTRUCT_fg and their BCg counterparts apply after each
// reconstruction (MainSrc/POD_ROUTINES.f08:28-353, MainSrc/DMD_ROUTINES.f08:131-400).
//
// A trt_romgrid registers the data-grid fields on its context (dat_fg_avg_xx ... dat_BCg);
// POD and DMD bases reconstruct into them.  trt_romgrid_map then, per field, appends the
// boundary values the reference appends (EXTND_BT_BND / EXTND_LR_BND for the face fields,
// AVG_APND_BND with COLLECT_*_BND for the cell fields, the side extension of BCg) and
// interpolates onto the simulation grid (FMAP, bilinear through the MAP_GRIDS maps),
// writing the *_next arrays of trt_step_given_qdf.
//
// MAP_GRIDS is a brute-force nearest-point search (first minimum in point order) followed
// by the choice of the enclosing cell's four points; it runs once, at creation, as one
// GPU thread per simulation point over all data points, with the reference's distance
// (sqrt of the sum of squares, correctly rounded) and 1e-10 tolerances.
#include "trt_internal.h"

#include <cmath>

struct trt_romgrid {
    trt_ctx *ctx = nullptr;
    int dnx = 0, dny = 0, ng = 0, nx = 0, ny = 0, dnt = 0, nt = 0;
    double *grid = nullptr;      // all point grids, device
    int *gmap = nullptr;         // all maps, device: Map (1-based coordinate positions), 4 per point
    int *vmap = nullptr;         // (Map + 1) / 2: 1-based point indices
    double *f2 = nullptr;        // augmented data field (largest of the layouts x N_g)
    // offsets into grid / maps
    size_t g_sim[7], g_dat[7], m_off[7];
    int npts_sim[7], npts_dat[7];
    std::vector<int> tmap;       // 2 per simulation time, 1-based data times
    std::vector<double> sim_t, dat_t;
    bool square = false;
};

namespace {

// layouts: 0 cell centres, 1 V faces, 2 H faces, 3..6 left, bottom, right, top sides
enum { L_AVG = 0, L_EDGV = 1, L_EDGH = 2, L_BND = 3 };

void grid_avg(std::vector<double> &o, const std::vector<double> &dx, const std::vector<double> &dy)
{
    const int nx = (int)dx.size(), ny = (int)dy.size();
    std::vector<double> xl(nx);
    xl[0] = dx[0] / 2.0;
    for (int i = 1; i < nx; ++i) xl[i] = xl[i - 1] + (dx[i - 1] + dx[i]) / 2.0;
    double y = 0.0;
    for (int j = 0; j < ny; ++j) {
        y = y + dy[j] / 2.0;
        for (int i = 0; i < nx; ++i) { o.push_back(xl[i]); o.push_back(y); }
        y = y + dy[j] / 2.0;
    }
}

void grid_edgv(std::vector<double> &o, const std::vector<double> &dx, const std::vector<double> &dy, int nxp)
{
    std::vector<double> xl(nxp);
    xl[0] = 0.0;
    for (int i = 1; i < nxp; ++i) xl[i] = xl[i - 1] + dx[i - 1];
    double y = 0.0;
    for (size_t j = 0; j < dy.size(); ++j) {
        y = y + dy[j] / 2.0;
        for (int i = 0; i < nxp; ++i) { o.push_back(xl[i]); o.push_back(y); }
        y = y + dy[j] / 2.0;
    }
}

void grid_edgh(std::vector<double> &o, const std::vector<double> &dx, const std::vector<double> &dy, int nyp)
{
    const int nx = (int)dx.size();
    std::vector<double> xl(nx);
    xl[0] = dx[0] / 2.0;
    for (int i = 1; i < nx; ++i) xl[i] = xl[i - 1] + (dx[i - 1] + dx[i]) / 2.0;
    double y = 0.0;
    for (int j = 0; j < nyp; ++j) {
        for (int i = 0; i < nx; ++i) { o.push_back(xl[i]); o.push_back(y); }
        if (j < (int)dy.size()) y = y + dy[j];
    }
}

void grid_lr(std::vector<double> &o, double x, const std::vector<double> &dy)
{
    double y = 0.0;
    for (double d : dy) {
        y = y + d / 2.0;
        o.push_back(x);
        o.push_back(y);
        y = y + d / 2.0;
    }
}

void grid_bt(std::vector<double> &o, const std::vector<double> &dx, double y)
{
    double x = 0.0;
    for (double d : dx) {
        x = x + d / 2.0;
        o.push_back(x);
        o.push_back(y);
        x = x + d / 2.0;
    }
}

// MAP_GRIDS (GRID_FUNCTIONS.f08:222-327): one thread per grid2 point; npts1 grid1 points
// are searched (positions beyond may be neighbours)
__global__ void map_grids_kernel(const double *__restrict__ g1, int npts1, const double *__restrict__ g2, int npts2,
                                 int xlen1, int *__restrict__ gmap, int *__restrict__ vmap)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= npts2) return;
    const double x2 = g2[2 * j], y2 = g2[2 * j + 1];
    double d = HUGE_VAL;
    int l1 = 1;
    for (int i = 0; i < npts1; ++i) {
        const double a = g1[2 * i] - x2, b = g1[2 * i + 1] - y2;
        const double r = fabs(sqrt(a * a + b * b));
        if (r < d) { d = r; l1 = 2 * i + 1; }
    }
    const double gx = g1[l1 - 1], gy = g1[l1];
    const double tol = 1e-10;
    int l[4];
    if (gx - x2 < -tol) {
        if (gy - y2 < -tol) { l[0] = l1; l[1] = l1 + 2; l[2] = l1 + xlen1; l[3] = l1 + xlen1 + 2; }
        else if (gy - y2 > tol) { const int a = l1 - xlen1; l[0] = a; l[1] = a + 2; l[2] = l1; l[3] = l1 + 2; }
        else { l[0] = l1; l[1] = l1 + 2; l[2] = l1; l[3] = l1 + 2; }
    } else if (gx - x2 > tol) {
        if (gy - y2 < -tol) { const int a = l1 - 2; l[0] = a; l[1] = l1; l[2] = a + xlen1; l[3] = a + xlen1 + 2; }
        else if (gy - y2 > tol) { const int b = l1 - xlen1; l[0] = b - 2; l[1] = b; l[2] = l1 - 2; l[3] = l1; }
        else { const int a = l1 - 2; l[0] = a; l[1] = l1; l[2] = a; l[3] = l1; }
    } else {
        if (gy - y2 < -tol) { l[0] = l1; l[1] = l1; l[2] = l1 + xlen1; l[3] = l1 + xlen1; }
        else if (gy - y2 > tol) { const int a = l1 - xlen1; l[0] = a; l[1] = a; l[2] = l1; l[3] = l1; }
        else { l[0] = l[1] = l[2] = l[3] = l1; }
    }
    for (int k = 0; k < 4; ++k) {
        gmap[4 * j + k] = l[k];
        vmap[4 * j + k] = (l[k] + 1) / 2;
    }
}

// BILINEAR_INTERPOLATE (GRID_FUNCTIONS.f08:366-400)
__device__ __forceinline__ double bilinear(double f11, double f12, double f21, double f22, double x, double x1,
                                           double x2, double y, double y1, double y2)
{
    if ((x2 - x1) == 0.0) {
        if ((y2 - y1) == 0.0) return f11;
        const double yy1 = y - y1, yy2 = y2 - y;
        return (f11 * yy2 + f12 * yy1) / (y2 - y1);
    }
    if ((y2 - y1) == 0.0) {
        const double xx1 = x - x1, xx2 = x2 - x;
        return (f11 * xx2 + f21 * xx1) / (x2 - x1);
    }
    const double xx1 = x - x1, xx2 = x2 - x, yy1 = y - y1, yy2 = y2 - y;
    return (f11 * xx2 * yy2 + f21 * xx1 * yy2 + f12 * xx2 * yy1 + f22 * xx1 * yy1) / ((x2 - x1) * (y2 - y1));
}

// FMAP (GRID_FUNCTIONS.f08:14-43): out[g glen2 + j] from f1[g glen1 + ...]
__global__ void fmap_kernel(const double *__restrict__ f1, const double *__restrict__ g1,
                            const double *__restrict__ g2, const int *__restrict__ gm, const int *__restrict__ vm,
                            int glen1, int glen2, int ng, double *__restrict__ out)
{
    const size_t n = (size_t)glen2 * ng;
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
        const int g = (int)(k / glen2), j = (int)(k - (size_t)g * glen2);
        const double *f = f1 + (size_t)g * glen1 - 1;   // 1-based point indices
        const int *v = vm + 4 * j, *m = gm + 4 * j;
        out[k] = bilinear(f[v[0]], f[v[2]], f[v[1]], f[v[3]], g2[2 * j], g1[m[0] - 1], g1[m[3] - 1], g2[2 * j + 1],
                          g1[m[0]], g1[m[3]]);
    }
}

// The boundary values appended before the map (GRID_FUNCTIONS.f08:406-659 as
// POD_RECONSTRUCT_fg combines them, POD_ROUTINES.f08:81-232), one output element per
// thread.  kind 0: EXTND_BT_BND of a V-face field (dN_x+1, dN_y) -> (dN_x+1, dN_y+2);
// 1: EXTND_LR_BND of an H-face field (dN_x, dN_y+1) -> (dN_x+2, dN_y+1);
// 2: AVG_APND_BND of fg_avg_xx: left/right from fg_edgV_xx's boundary faces
// (COLLECT_LR_BND), bottom/top rows and corners from the field itself;
// 3: AVG_APND_BND of fg_avg_yy: bottom/top from fg_edgH_yy's boundary faces
// (COLLECT_BT_BND), left/right columns and corners from the field itself;
// 4: one side of BCg (n values per group at offset off of 2(dN_x+dN_y)) extended by its
// end values -> n + 2 per group (POD_ROUTINES.f08:285-323).
__global__ void augment_kernel(int kind, const double *__restrict__ f, const double *__restrict__ side,
                               int dnx, int dny, int ng, int off, int n, double *__restrict__ out)
{
    size_t per;
    switch (kind) {
    case 0: per = (size_t)(dnx + 1) * (dny + 2); break;
    case 1: per = (size_t)(dnx + 2) * (dny + 1); break;
    case 4: per = (size_t)(n + 2); break;
    default: per = (size_t)(dnx + 2) * (dny + 2);
    }
    const size_t tot = per * ng;
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < tot; k += (size_t)gridDim.x * blockDim.x) {
        const int g = (int)(k / per);
        const int p = (int)(k - (size_t)g * per);
        double v;
        if (kind == 0) {
            const int w = dnx + 1, io = p % w, jo = p / w;
            const int j = min(max(jo - 1, 0), dny - 1);
            v = f[io + (size_t)w * (j + (size_t)dny * g)];
        } else if (kind == 1) {
            const int w = dnx + 2, io = p % w, jo = p / w;
            const int i = min(max(io - 1, 0), dnx - 1);
            v = f[i + (size_t)dnx * (jo + (size_t)(dny + 1) * g)];
        } else if (kind == 4) {
            const int i = min(max(p - 1, 0), n - 1);
            v = f[(size_t)2 * (dnx + dny) * g + off + i];
        } else {
            const int w = dnx + 2, io = p % w, jo = p / w;
            auto at = [&](int i, int j) { return f[i + (size_t)dnx * (j + (size_t)dny * g)]; };
            const bool left = io == 0, right = io == dnx + 1, bot = jo == 0, top = jo == dny + 1;
            if ((left || right) && (bot || top)) {
                v = at(left ? 0 : dnx - 1, bot ? 0 : dny - 1);                      // COLLECT_CRN_BND
            } else if (left || right) {
                if (kind == 2)                                                       // fg_edgV_xx faces
                    v = side[(left ? 0 : dnx) + (size_t)(dnx + 1) * ((jo - 1) + (size_t)dny * g)];
                else
                    v = at(left ? 0 : dnx - 1, jo - 1);
            } else if (bot || top) {
                if (kind == 3)                                                       // fg_edgH_yy faces
                    v = side[(io - 1) + (size_t)dnx * ((bot ? 0 : dny) + (size_t)(dny + 1) * g)];
                else
                    v = at(io - 1, bot ? 0 : dny - 1);
            } else {
                v = at(io - 1, jo - 1);
            }
        }
        out[k] = v;
    }
}

int blocks_for(size_t n) { return (int)std::min<size_t>((n + 255) / 256, 16384); }

template <typename F>
int rguarded(trt_ctx *ctx, F &&f)
{
    if (!ctx) return -1;
    try {
        f(ctx->c);
        return 0;
    } catch (const trt::Error &e) {
        ctx->c.err = e.msg;
    } catch (const std::exception &e) {
        ctx->c.err = e.what();
    }
    return -2;
}

void romgrid_free(trt_romgrid *r)
{
    for (void *p : {(void *)r->grid, (void *)r->gmap, (void *)r->vmap, (void *)r->f2})
        if (p) (void)hipFree(p);
    delete r;
}

const char *const kFields[6] = {"fg_avg_xx", "fg_avg_yy", "fg_edgV_xx", "fg_edgV_xy", "fg_edgH_yy", "fg_edgH_xy"};

}  // namespace

namespace trt {

void romgrid_release_all(Ctx &c)
{
    for (trt_romgrid *r : c.romgrids) romgrid_free(r);
    c.romgrids.clear();
}

}  // namespace trt

extern "C" {

int trt_romgrid_create(trt_ctx *ctx, trt_romgrid **out, int dnx, int dny, const double *delx_d, const double *dely_d,
                       double xlen_d, double ylen_d, int dnt, double delt_d, double start_time_d, int nt,
                       double start_time, double xlen, double ylen)
{
    return rguarded(ctx, [&](trt::Ctx &c) {
        if (!out || !delx_d || !dely_d) throw trt::Error("null argument");
        *out = nullptr;
        if (dnx < 2 || dny < 2) throw trt::Error("the data grid needs at least 2 x 2 cells");
        if (dnt < 1 || nt < 1) throw trt::Error("need at least one data time and one simulation time");
        for (const char *f : kFields) {
            const std::string n = std::string("dat_") + f;
            if (c.arrays.count(n)) throw trt::Error("the context already has a data grid");
        }
        trt_romgrid *r = new trt_romgrid;
        try {
            r->ctx = ctx;
            r->dnx = dnx; r->dny = dny; r->ng = c.ng; r->nx = c.nx; r->ny = c.ny; r->dnt = dnt; r->nt = nt;
            r->square = dnx == dny && c.nx == c.ny;
            std::vector<double> dxd(delx_d, delx_d + dnx), dyd(dely_d, dely_d + dny);
            std::vector<double> dx(c.nx), dy(c.ny);
            TRT_CHECK(hipMemcpy(dx.data(), c.dx, c.nx * sizeof(double), hipMemcpyDeviceToHost));
            TRT_CHECK(hipMemcpy(dy.data(), c.dy, c.ny * sizeof(double), hipMemcpyDeviceToHost));
            std::vector<double> dx2(dnx + 2, 0.0), dy2(dny + 2, 0.0);
            for (int i = 0; i < dnx; ++i) dx2[i + 1] = dxd[i];
            for (int j = 0; j < dny; ++j) dy2[j + 1] = dyd[j];
            // GENERATE_GRIDS_XY (GRID_FUNCTIONS.f08:689-752); the right side at x = ylen and the top
            // at y = xlen, as the reference passes them
            std::vector<double> g;
            auto mark = [&](size_t *o, int k) { o[k] = g.size(); };
            mark(r->g_sim, L_AVG); grid_avg(g, dx, dy);
            mark(r->g_sim, L_EDGV); grid_edgv(g, dx, dy, c.nx + 1);
            mark(r->g_sim, L_EDGH); grid_edgh(g, dx, dy, c.ny + 1);
            mark(r->g_sim, 3); grid_lr(g, 0.0, dy);
            mark(r->g_sim, 4); grid_bt(g, dx, 0.0);
            mark(r->g_sim, 5); grid_lr(g, ylen, dy);
            mark(r->g_sim, 6); grid_bt(g, dx, xlen);
            mark(r->g_dat, L_AVG); grid_avg(g, dx2, dy2);
            mark(r->g_dat, L_EDGV); grid_edgv(g, dxd, dy2, dnx + 1);
            mark(r->g_dat, L_EDGH); grid_edgh(g, dx2, dyd, dny + 1);
            mark(r->g_dat, 3); grid_lr(g, 0.0, dy2);
            mark(r->g_dat, 4); grid_bt(g, dx2, 0.0);
            mark(r->g_dat, 5); grid_lr(g, ylen_d, dy2);
            mark(r->g_dat, 6); grid_bt(g, dx2, xlen_d);
            const int ns[7] = {c.nx * c.ny, (c.nx + 1) * c.ny, c.nx * (c.ny + 1), c.ny, c.nx, c.ny, c.nx};
            const int nd[7] = {(dnx + 2) * (dny + 2), (dnx + 1) * (dny + 2), (dnx + 2) * (dny + 1), dny + 2, dnx + 2,
                               dny + 2, dnx + 2};
            // MAP_GRIDS arguments (:743-750): search lengths and the row widths in coordinates;
            // every side searches dN_y + 1 points and maps N_y points
            const int search[7] = {nd[0], nd[1], nd[2], dny + 1, dny + 1, dny + 1, dny + 1};
            const int mapped[7] = {ns[0], ns[1], ns[2], c.ny, c.ny, c.ny, c.ny};
            const int xl1[7] = {2 * (dnx + 2), 2 * (dnx + 1), 2 * (dnx + 2), 2, 0, 2, 0};
            size_t mo = 0;
            for (int k = 0; k < 7; ++k) {
                r->npts_sim[k] = ns[k];
                r->npts_dat[k] = nd[k];
                r->m_off[k] = mo;
                mo += (size_t)4 * std::max(ns[k], mapped[k]);
            }
            TRT_CHECK(hipMalloc(reinterpret_cast<void **>(&r->grid), g.size() * sizeof(double)));
            TRT_CHECK(hipMemcpy(r->grid, g.data(), g.size() * sizeof(double), hipMemcpyHostToDevice));
            TRT_CHECK(hipMalloc(reinterpret_cast<void **>(&r->gmap), mo * sizeof(int)));
            TRT_CHECK(hipMalloc(reinterpret_cast<void **>(&r->vmap), mo * sizeof(int)));
            TRT_CHECK(hipMemsetAsync(r->gmap, 0, mo * sizeof(int), c.stream));
            TRT_CHECK(hipMemsetAsync(r->vmap, 0, mo * sizeof(int), c.stream));
            for (int k = 0; k < 7; ++k) {
                if (k >= 3 && !r->square) continue;   // the sides' maps cover them only on square grids
                hipLaunchKernelGGL(map_grids_kernel, dim3((mapped[k] + 127) / 128), dim3(128), 0, c.stream,
                                   r->grid + r->g_dat[k], search[k], r->grid + r->g_sim[k], mapped[k], xl1[k],
                                   r->gmap + r->m_off[k], r->vmap + r->m_off[k]);
                TRT_CHECK(hipGetLastError());
            }
            // every neighbour must be a point of the data grid: a simulation point outside the data
            // domain would reach beyond it (the reference then reads past its arrays)
            {
                std::vector<int> hm(mo);
                TRT_CHECK(hipMemcpyAsync(hm.data(), r->gmap, mo * sizeof(int), hipMemcpyDeviceToHost, c.stream));
                TRT_CHECK(hipStreamSynchronize(c.stream));
                for (int k = 0; k < 7; ++k) {
                    if (k >= 3 && !r->square) continue;
                    for (size_t e = 0; e < (size_t)4 * mapped[k]; ++e) {
                        const int l = hm[r->m_off[k] + e];
                        if (l < 1 || l > 2 * nd[k] - 1 || l % 2 == 0)
                            throw trt::Error("a simulation point lies outside the data grid (layout " + std::to_string(k) +
                                             ", point " + std::to_string(e / 4 + 1) + ")");
                    }
                }
            }
            size_t fmax = 0;
            for (int k = 0; k < 7; ++k) fmax = std::max(fmax, (size_t)nd[k]);
            TRT_CHECK(hipMalloc(reinterpret_cast<void **>(&r->f2), fmax * c.ng * sizeof(double)));
            // GENERATE_GRIDS_T (:754-817)
            r->dat_t.resize(dnt);
            r->sim_t.resize(nt);
            r->dat_t[0] = start_time_d + delt_d;
            for (int j = 1; j < dnt; ++j) r->dat_t[j] = r->dat_t[j - 1] + delt_d;
            r->sim_t[0] = start_time + c.p.dt;
            for (int j = 1; j < nt; ++j) r->sim_t[j] = r->sim_t[j - 1] + c.p.dt;
            for (int i = 0; i < nt; ++i) {
                const double s = r->sim_t[i];
                int a = 0, b = 0;
                if (r->dat_t[0] >= s) a = b = 1;
                else if (r->dat_t[dnt - 1] <= s) a = b = dnt;
                else
                    for (int j = 2; j <= dnt; ++j) {
                        if (std::fabs(r->dat_t[j - 1] - s) / r->dat_t[j - 1] <= 1e-14) { a = b = j; break; }
                        if (r->dat_t[j - 1] > s) { a = j; b = j - 1; break; }
                    }
                r->tmap.push_back(a);
                r->tmap.push_back(b);
            }
            // the data-grid fields the bases reconstruct into
            c.reg("dat_fg_avg_xx", dnx, dny, c.ng);
            c.reg("dat_fg_avg_yy", dnx, dny, c.ng);
            c.reg("dat_fg_edgV_xx", dnx + 1, dny, c.ng);
            c.reg("dat_fg_edgV_xy", dnx + 1, dny, c.ng);
            c.reg("dat_fg_edgH_yy", dnx, dny + 1, c.ng);
            c.reg("dat_fg_edgH_xy", dnx, dny + 1, c.ng);
            c.reg("dat_BCg", 2 * (dnx + dny), c.ng);
            TRT_CHECK(hipStreamSynchronize(c.stream));
        } catch (...) {
            romgrid_free(r);
            throw;
        }
        c.romgrids.push_back(r);
        *out = r;
    });
}

int trt_romgrid_tmap(const trt_romgrid *r, int t, int *t1, int *t2, double *x, double *x1, double *x2)
{
    if (!r || !t1 || !t2 || !x || !x1 || !x2) return -1;
    if (t < 1 || t > r->nt) {
        r->ctx->c.err = "simulation time index out of range";
        return -2;
    }
    *t1 = r->tmap[2 * (t - 1)];
    *t2 = r->tmap[2 * (t - 1) + 1];
    *x = r->sim_t[t - 1];
    *x1 = r->dat_t[*t1 - 1];
    *x2 = r->dat_t[*t2 - 1];
    return 0;
}

int trt_romgrid_map(trt_romgrid *r, int with_bcg)
{
    if (!r) return -1;
    return rguarded(r->ctx, [&](trt::Ctx &c) {
        const int dnx = r->dnx, dny = r->dny, ng = r->ng;
        auto fm = [&](int layout, const char *dst) {
            const size_t n = (size_t)r->npts_sim[layout] * ng;
            hipLaunchKernelGGL(fmap_kernel, dim3(blocks_for(n)), dim3(256), 0, c.stream, r->f2, r->grid + r->g_dat[layout],
                               r->grid + r->g_sim[layout], r->gmap + r->m_off[layout], r->vmap + r->m_off[layout],
                               r->npts_dat[layout], r->npts_sim[layout], ng, c.arr(dst));
            TRT_CHECK(hipGetLastError());
        };
        auto aug = [&](int kind, const char *src, const char *side, int off, int n, size_t len) {
            hipLaunchKernelGGL(augment_kernel, dim3(blocks_for(len * ng)), dim3(256), 0, c.stream, kind, c.arr(src),
                               side ? c.arr(side) : nullptr, dnx, dny, ng, off, n, r->f2);
            TRT_CHECK(hipGetLastError());
        };
        const size_t lv = (size_t)(dnx + 1) * (dny + 2), lh = (size_t)(dnx + 2) * (dny + 1),
                     la = (size_t)(dnx + 2) * (dny + 2);
        aug(0, "dat_fg_edgV_xx", nullptr, 0, 0, lv); fm(L_EDGV, "fg_edgV_xx_next");
        aug(0, "dat_fg_edgV_xy", nullptr, 0, 0, lv); fm(L_EDGV, "fg_edgV_xy_next");
        aug(1, "dat_fg_edgH_yy", nullptr, 0, 0, lh); fm(L_EDGH, "fg_edgH_yy_next");
        aug(1, "dat_fg_edgH_xy", nullptr, 0, 0, lh); fm(L_EDGH, "fg_edgH_xy_next");
        aug(2, "dat_fg_avg_xx", "dat_fg_edgV_xx", 0, 0, la); fm(L_AVG, "fg_avg_xx_next");
        aug(3, "dat_fg_avg_yy", "dat_fg_edgH_yy", 0, 0, la); fm(L_AVG, "fg_avg_yy_next");
        if (with_bcg) {
            if (!r->square)
                throw trt::Error("C_g's boundary map needs N_x = N_y and dN_x = dN_y (the reference maps every "
                                 "side with N_y and dN_y points, GRID_FUNCTIONS.f08:743-750)");
            const int off[4] = {0, dny, dny + dnx, 2 * dny + dnx}, n[4] = {dny, dnx, dny, dnx};
            const char *dst[4] = {"Cg_L_next", "Cg_B_next", "Cg_R_next", "Cg_T_next"};
            for (int s = 0; s < 4; ++s) {
                aug(4, "dat_BCg", nullptr, off[s], n[s], (size_t)n[s] + 2);
                fm(3 + s, dst[s]);
            }
        }
        TRT_CHECK(hipStreamSynchronize(c.stream));
    });
}

int trt_romgrid_maps(const trt_romgrid *r, int layout, int *gmap, int *vmap, double *sim_grid, double *dat_grid)
{
    if (!r || layout < 0 || layout > 6) return -1;
    const size_t n = (size_t)4 * r->npts_sim[layout];
    if (gmap) (void)hipMemcpy(gmap, r->gmap + r->m_off[layout], n * sizeof(int), hipMemcpyDeviceToHost);
    if (vmap) (void)hipMemcpy(vmap, r->vmap + r->m_off[layout], n * sizeof(int), hipMemcpyDeviceToHost);
    if (sim_grid)
        (void)hipMemcpy(sim_grid, r->grid + r->g_sim[layout], 2 * r->npts_sim[layout] * sizeof(double),
                        hipMemcpyDeviceToHost);
    if (dat_grid)
        (void)hipMemcpy(dat_grid, r->grid + r->g_dat[layout], 2 * r->npts_dat[layout] * sizeof(double),
                        hipMemcpyDeviceToHost);
    return 0;
}

int trt_romgrid_destroy(trt_romgrid *r)
{
    if (!r) return -1;
    trt::Ctx &c = r->ctx->c;
    (void)hipStreamSynchronize(c.stream);
    for (size_t k = 0; k < c.romgrids.size(); ++k)
        if (c.romgrids[k] == r) {
            c.romgrids.erase(c.romgrids.begin() + k);
            break;
        }
    romgrid_free(r);
    return 0;
}

}  // extern "C"
