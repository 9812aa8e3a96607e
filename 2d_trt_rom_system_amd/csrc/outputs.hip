// The reference's NetCDF output file, written from a context's device state:
// OUTFILE_INIT (MainSrc/OUTPUTS.f08:13-408), OUTFILE_INIT_Tgen (:412-835),
// OUTFILE_VARDEFS (:839-1425), TIMESTEP_OUTS (:1429-1620), the per-step writes of
// TRT_MLQD_ALGORITHM (MainSrc/ALGORITHMS.f08:688-868: iteration counts, convergence
// norms, wave speeds) and of Tgen_QDf (:1036-1095); the wave speeds of
// MainSrc/WAVEPROP_TOOLS.f08.  Same dimension, variable and attribute names, types,
// units and shapes; the container is the classic format (ncfile.hip) instead of
// NETCDF4/HDF5 (NCDF_IO.f08:71), which netCDF readers open the same way.  Host code.
#include "ncfile.h"
#include "trt_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>

using trt::Error;
namespace nc = trt::nc;

namespace {

enum Flag { F_I, F_HOEG, F_HOFG, F_HOE, F_HOF, F_EG, F_FG, F_MGE, F_MGF, F_QDFG, F_E, F_F, F_D, F_OLD, F_KAP, F_SRC,
            F_ALWAYS };

// One variable of the file: its name, the context array it comes from, its Fortran
// dimensions (x = N_x, y = N_y, m = N_m, g = N_g, t = N_t, V = N_edgV, H = N_edgH), units,
// grid attributes (c: cell centres, v: V faces, h: H faces, b: tpts only), the output
// flag that defines it, and whether TIMESTEP_OUTS writes it only with use_grey = 1.
struct OutVar {
    const char *name, *arr, *dims, *units;
    char grid;
    int flag;
    bool grey;
};

constexpr const char *EU = "ev/cm^3";
constexpr const char *IU = "erg/(Ster*cm^3)";

// OUTFILE_VARDEFS, in its order (OUTPUTS.f08:925-1366; the iteration counts and norms of
// :901-921 are defined when the file is closed, see trt_outfile_close)
const OutVar kRunVars[] = {
    {"Cb_L", "Cb_L", "y,t", nullptr, 0, F_ALWAYS, true}, {"Cb_B", "Cb_B", "x,t", nullptr, 0, F_ALWAYS, true},
    {"Cb_R", "Cb_R", "y,t", nullptr, 0, F_ALWAYS, true}, {"Cb_T", "Cb_T", "x,t", nullptr, 0, F_ALWAYS, true},
    {"E_in_L", "E_in_L", "y,t", nullptr, 0, F_ALWAYS, true}, {"E_in_B", "E_in_B", "x,t", nullptr, 0, F_ALWAYS, true},
    {"E_in_R", "E_in_R", "y,t", nullptr, 0, F_ALWAYS, true}, {"E_in_T", "E_in_T", "x,t", nullptr, 0, F_ALWAYS, true},
    {"F_in_L", "F_in_L", "y,t", nullptr, 0, F_ALWAYS, true}, {"F_in_B", "F_in_B", "x,t", nullptr, 0, F_ALWAYS, true},
    {"F_in_R", "F_in_R", "y,t", nullptr, 0, F_ALWAYS, true}, {"F_in_T", "F_in_T", "x,t", nullptr, 0, F_ALWAYS, true},
    {"Cg_L", "Cg_L", "y,g,t", nullptr, 'b', F_ALWAYS, false}, {"Cg_B", "Cg_B", "x,g,t", nullptr, 'b', F_ALWAYS, false},
    {"Cg_R", "Cg_R", "y,g,t", nullptr, 'b', F_ALWAYS, false}, {"Cg_T", "Cg_T", "x,g,t", nullptr, 'b', F_ALWAYS, false},
    {"Eg_in_L", "Eg_in_L", "y,g,t", nullptr, 0, F_ALWAYS, false},
    {"Eg_in_B", "Eg_in_B", "x,g,t", nullptr, 0, F_ALWAYS, false},
    {"Eg_in_R", "Eg_in_R", "y,g,t", nullptr, 0, F_ALWAYS, false},
    {"Eg_in_T", "Eg_in_T", "x,g,t", nullptr, 0, F_ALWAYS, false},
    {"Fg_in_L", "Fg_in_L", "y,g,t", nullptr, 0, F_ALWAYS, false},
    {"Fg_in_B", "Fg_in_B", "x,g,t", nullptr, 0, F_ALWAYS, false},
    {"Fg_in_R", "Fg_in_R", "y,g,t", nullptr, 0, F_ALWAYS, false},
    {"Fg_in_T", "Fg_in_T", "x,g,t", nullptr, 0, F_ALWAYS, false},
    {"Temp_xWave_Speed", nullptr, "t", nullptr, 0, F_ALWAYS, false},
    {"E_xWave_Speed", nullptr, "t", nullptr, 0, F_ALWAYS, false},
    {"Temperature", "Temp", "x,y,t", EU, 0, F_ALWAYS, false},
    {"E_avg_Grey", "E_avg", "x,y,t", EU, 'c', F_E, true},
    {"E_edgV_Grey", "E_edgV", "V,y,t", EU, 'v', F_E, true},
    {"E_edgH_Grey", "E_edgH", "x,H,t", EU, 'h', F_E, true},
    {"E_avg_MGQD", "MGQD_E_avg", "x,y,t", EU, 'c', F_MGE, false},
    {"E_edgV_MGQD", "MGQD_E_edgV", "V,y,t", EU, 'v', F_MGE, false},
    {"E_edgH_MGQD", "MGQD_E_edgH", "x,H,t", EU, 'h', F_MGE, false},
    {"E_avg_HO", "HO_E_avg", "x,y,t", EU, 'c', F_HOE, false},
    {"E_edgV_HO", "HO_E_edgV", "V,y,t", EU, 'v', F_HOE, false},
    {"E_edgH_HO", "HO_E_edgH", "x,H,t", EU, 'h', F_HOE, false},
    {"Fx_edgV_Grey", "Fx_edgV", "V,y,t", EU, 'v', F_F, true},
    {"Fy_edgH_Grey", "Fy_edgH", "x,H,t", EU, 'h', F_F, true},
    {"Fx_edgV_MGQD", "MGQD_Fx_edgV", "V,y,t", EU, 'v', F_MGF, false},
    {"Fy_edgH_MGQD", "MGQD_Fy_edgH", "x,H,t", EU, 'h', F_MGF, false},
    {"Fx_edgV_HO", "HO_Fx_edgV", "V,y,t", EU, 'v', F_HOF, false},
    {"Fy_edgH_HO", "HO_Fy_edgH", "x,H,t", EU, 'h', F_HOF, false},
    {"Eg_avg_MGQD", "Eg_avg", "x,y,g,t", EU, 'c', F_EG, false},
    {"Eg_edgV_MGQD", "Eg_edgV", "V,y,g,t", EU, 'v', F_EG, false},
    {"Eg_edgH_MGQD", "Eg_edgH", "x,H,g,t", EU, 'h', F_EG, false},
    {"Eg_avg_HO", "HO_Eg_avg", "x,y,g,t", EU, 'c', F_HOEG, false},
    {"Eg_edgV_HO", "HO_Eg_edgV", "V,y,g,t", EU, 'v', F_HOEG, false},
    {"Eg_edgH_HO", "HO_Eg_edgH", "x,H,g,t", EU, 'h', F_HOEG, false},
    {"Fxg_edgV_MGQD", "Fxg_edgV", "V,y,g,t", EU, 'v', F_FG, false},
    {"Fyg_edgH_MGQD", "Fyg_edgH", "x,H,g,t", EU, 'h', F_FG, false},
    {"Fxg_edgV_HO", "HO_Fxg_edgV", "V,y,g,t", EU, 'v', F_HOFG, false},
    {"Fyg_edgH_HO", "HO_Fyg_edgH", "x,H,g,t", EU, 'h', F_HOFG, false},
    {"I_avg", "I_avg", "x,y,m,g,t", IU, 'c', F_I, false},
    {"I_edgV", "I_edgV", "V,y,m,g,t", IU, 'v', F_I, false},
    {"I_edgH", "I_edgH", "x,H,m,g,t", IU, 'h', F_I, false},
    // the reference gives KapE and KapR no Units: its NF_DEF_UNIT calls for them name KapB
    // (OUTPUTS.f08:1277-1287)
    {"KapE_Bar", "KapE_Bar", "x,y,t", "1/cm", 0, F_KAP, false},
    {"KapB", "KapB", "x,y,g,t", "1/cm", 0, F_KAP, false},
    {"KapE", "KapE", "x,y,g,t", nullptr, 0, F_KAP, false},
    {"KapR", "KapR", "x,y,g,t", nullptr, 0, F_KAP, false},
    {"Bg", "Bg", "x,y,g,t", nullptr, 0, F_SRC, false},
    {"fg_avg_xx", "fg_avg_xx", "x,y,g,t", nullptr, 'c', F_QDFG, false},
    {"fg_avg_yy", "fg_avg_yy", "x,y,g,t", nullptr, 'c', F_QDFG, false},
    {"fg_avg_xy", "fg_avg_xy", "x,y,g,t", nullptr, 'c', F_QDFG, false},
    {"fg_edgV_xx", "fg_edgV_xx", "V,y,g,t", nullptr, 'v', F_QDFG, false},
    {"fg_edgV_xy", "fg_edgV_xy", "V,y,g,t", nullptr, 'v', F_QDFG, false},
    {"fg_edgH_yy", "fg_edgH_yy", "x,H,g,t", nullptr, 'h', F_QDFG, false},
    {"fg_edgH_xy", "fg_edgH_xy", "x,H,g,t", nullptr, 'h', F_QDFG, false},
    {"DC_xx", "DC_xx", "x,y,t", nullptr, 0, F_D, true}, {"DL_xx", "DL_xx", "x,y,t", nullptr, 0, F_D, true},
    {"DR_xx", "DR_xx", "x,y,t", nullptr, 0, F_D, true}, {"DC_yy", "DC_yy", "x,y,t", nullptr, 0, F_D, true},
    {"DB_yy", "DB_yy", "x,y,t", nullptr, 0, F_D, true}, {"DT_yy", "DT_yy", "x,y,t", nullptr, 0, F_D, true},
    {"DL_xy", "DL_xy", "x,y,t", nullptr, 0, F_D, true}, {"DB_xy", "DB_xy", "x,y,t", nullptr, 0, F_D, true},
    {"DR_xy", "DR_xy", "x,y,t", nullptr, 0, F_D, true}, {"DT_xy", "DT_xy", "x,y,t", nullptr, 0, F_D, true},
    {"Gold_L", "G_old", "x,y,g,t", nullptr, 0, F_OLD, false},
    {"Pold_L", "Pold_L", "x,y,g,t", nullptr, 0, F_OLD, false},
    {"Pold_B", "Pold_B", "x,y,g,t", nullptr, 0, F_OLD, false},
    {"Pold_R", "Pold_R", "x,y,g,t", nullptr, 0, F_OLD, false},
    {"Pold_T", "Pold_T", "x,y,g,t", nullptr, 0, F_OLD, false},
    {"Gold_hat", "Gold_Hat", "x,y,t", nullptr, 0, F_OLD, true},
    {"Rhat_old", "Rhat_old", "x,y,t", nullptr, 0, F_OLD, true},
    {"PL", "PL", "x,y,t", nullptr, 0, F_OLD, true}, {"PB", "PB", "x,y,t", nullptr, 0, F_OLD, true},
    {"PR", "PR", "x,y,t", nullptr, 0, F_OLD, true}, {"PT", "PT", "x,y,t", nullptr, 0, F_OLD, true},
};

// OUTFILE_INIT_Tgen's variables, in its order (OUTPUTS.f08:535-715)
const OutVar kTgenVars[] = {
    {"Cg_L", "Cg_L", "y,g,t", nullptr, 'b', F_ALWAYS, false}, {"Cg_B", "Cg_B", "x,g,t", nullptr, 'b', F_ALWAYS, false},
    {"Cg_R", "Cg_R", "y,g,t", nullptr, 'b', F_ALWAYS, false}, {"Cg_T", "Cg_T", "x,g,t", nullptr, 'b', F_ALWAYS, false},
    {"Temperature", "Temp", "x,y,t", EU, 0, F_ALWAYS, false},
    {"E_avg_HO", "HO_E_avg", "x,y,t", EU, 'c', F_HOE, false},
    {"E_edgV_HO", "HO_E_edgV", "V,y,t", EU, 'v', F_HOE, false},
    {"E_edgH_HO", "HO_E_edgH", "x,H,t", EU, 'h', F_HOE, false},
    {"Fx_edgV_HO", "HO_Fx_edgV", "V,y,t", EU, 'v', F_HOF, false},
    {"Fy_edgH_HO", "HO_Fy_edgH", "x,H,t", EU, 'h', F_HOF, false},
    {"Eg_avg_HO", "HO_Eg_avg", "x,y,g,t", EU, 'c', F_HOEG, false},
    {"Eg_edgV_HO", "HO_Eg_edgV", "V,y,g,t", EU, 'v', F_HOEG, false},
    {"Eg_edgH_HO", "HO_Eg_edgH", "x,H,g,t", EU, 'h', F_HOEG, false},
    {"Fxg_edgV_HO", "HO_Fxg_edgV", "V,y,g,t", EU, 'v', F_HOFG, false},
    {"Fyg_edgH_HO", "HO_Fyg_edgH", "x,H,g,t", EU, 'h', F_HOFG, false},
    {"I_avg", "I_avg", "x,y,m,g,t", IU, 'c', F_I, false},
    {"I_edgV", "I_edgV", "V,y,m,g,t", IU, 'v', F_I, false},
    {"I_edgH", "I_edgH", "x,H,m,g,t", IU, 'h', F_I, false},
    {"fg_avg_xx", "fg_avg_xx", "x,y,g,t", nullptr, 'c', F_QDFG, false},
    {"fg_avg_yy", "fg_avg_yy", "x,y,g,t", nullptr, 'c', F_QDFG, false},
    {"fg_avg_xy", "fg_avg_xy", "x,y,g,t", nullptr, 'c', F_QDFG, false},
    {"fg_edgV_xx", "fg_edgV_xx", "V,y,g,t", nullptr, 'v', F_QDFG, false},
    {"fg_edgV_xy", "fg_edgV_xy", "V,y,g,t", nullptr, 'v', F_QDFG, false},
    {"fg_edgH_yy", "fg_edgH_yy", "x,H,g,t", nullptr, 'h', F_QDFG, false},
    {"fg_edgH_xy", "fg_edgH_xy", "x,H,g,t", nullptr, 'h', F_QDFG, false},
};

std::string text(const char *s)
{
    std::string r = s ? s : "";
    // Fortran CHARACTER(100) values arrive blank-padded; the attribute holds the word
    while (!r.empty() && r.back() == ' ') r.pop_back();
    return r;
}

// YMEAN, XWAVE_SPEED_edg and XWAVE_SPEED_cnt (MainSrc/WAVEPROP_TOOLS.f08:14-125): the
// y-average of column xp (1-based) of an array read with row stride N_x -- also for the
// N_x+1-long rows of E_edgV, as the reference reads them -- and the front speeds.
double ymean(const double *x, const double *dy, double ylen, int ny, int nx, int xp)
{
    double s = 0.0;
    size_t p = (size_t)xp - 1;
    for (int j = 0; j < ny; ++j) {
        s = s + x[p] * dy[j];
        p += (size_t)nx;
    }
    return s / ylen;
}

double xwave_speed_edg(const double *xc, const double *xe, const double *xold, const double *dy, const double *dx,
                       double dt, double ylen, int ny, int nx)
{
    double tsum = 0.0, xsum = 0.0;
    double xr = ymean(xe, dy, ylen, ny, nx, 1);
    for (int i = 1; i <= nx; ++i) {
        const double xl = xr;
        xr = ymean(xe, dy, ylen, ny, nx, i + 1);
        xsum = xsum + std::fabs((xr - xl) / dx[i - 1]);
        const double tl = ymean(xold, dy, ylen, ny, nx, i);
        const double tr = ymean(xc, dy, ylen, ny, nx, i);
        tsum = tsum + std::fabs((tr - tl) / dt);
    }
    return tsum / xsum;
}

double xwave_speed_cnt(const double *xc, const double *xold, const double *dy, const double *dx, double dt,
                       double ylen, int ny, int nx)
{
    double tsum = 0.0, xsum = 0.0;
    double xl = ymean(xc, dy, ylen, ny, nx, 1), xr = ymean(xc, dy, ylen, ny, nx, 2);
    xsum = xsum + std::fabs((xr - xl) / dx[0]);
    for (int i = 2; i <= nx - 1; ++i) {
        xl = ymean(xc, dy, ylen, ny, nx, i - 1);
        xr = ymean(xc, dy, ylen, ny, nx, i + 1);
        xsum = xsum + std::fabs((xr - xl) / dx[i - 1] / 2.0);
    }
    xl = ymean(xc, dy, ylen, ny, nx, nx - 1);
    xr = ymean(xc, dy, ylen, ny, nx, nx);
    xsum = xsum + std::fabs((xr - xl) / dx[nx - 1]);
    for (int i = 1; i <= nx; ++i) {
        const double tl = ymean(xold, dy, ylen, ny, nx, i);
        const double tr = ymean(xc, dy, ylen, ny, nx, i);
        tsum = tsum + std::fabs((tr - tl) / dt);
    }
    return tsum / xsum;
}

}  // namespace

struct trt_outfile {
    trt_ctx *ctx = nullptr;
    nc::Writer w;
    int kind = 0, nt = 0, use_grey = 1, its_out = 0, conv_out = 0;
    int flags[F_ALWAYS + 1] = {};
    bool fill = true;
    std::string err;
    std::vector<int> vid;                   // per table entry: variable id, -1 if not defined
    const OutVar *table = nullptr;
    size_t ntable = 0;
    std::map<int, std::vector<trt::RtItLog>> hist;   // iteration history per time index
    std::vector<double> dx, dy;
    double ylen = 0;
};

struct trt_ncread {
    nc::Reader r;
    std::string err;
};

namespace {

template <typename F>
int out_guarded(trt_outfile *f, F &&fn)
{
    if (!f) return -1;
    try {
        fn(*f);
        return 0;
    } catch (const Error &e) {
        f->err = e.msg;
    } catch (const std::exception &e) {
        f->err = e.what();
    }
    return -2;
}

std::vector<double> pull(trt_ctx *ctx, const char *name)
{
    const size_t n = trt_array_size(ctx, name);
    if (n == 0) throw Error(std::string("output: the context has no array ") + name);
    std::vector<double> v(n);
    if (trt_pull(ctx, name, v.data(), n) != 0) throw Error(std::string("output: ") + trt_last_error(ctx));
    return v;
}

// Fortran dimension list -> dimension ids, slowest first
std::vector<int> dims_of(const nc::Header &h, const char *spec)
{
    std::vector<int> f;
    for (const char *p = spec; *p; ++p) {
        const char *nm = nullptr;
        switch (*p) {
        case 'x': nm = "N_x"; break;
        case 'y': nm = "N_y"; break;
        case 'm': nm = "N_m"; break;
        case 'g': nm = "N_g"; break;
        case 't': nm = "N_t"; break;
        case 'V': nm = "N_edgV"; break;
        case 'H': nm = "N_edgH"; break;
        default: continue;
        }
        f.push_back(h.find_dim(nm));
    }
    std::reverse(f.begin(), f.end());
    return f;
}

void def_table(trt_outfile &o)
{
    o.vid.assign(o.ntable, -1);
    const int one = 1, three = 3;
    for (size_t k = 0; k < o.ntable; ++k) {
        const OutVar &v = o.table[k];
        if (v.flag != F_ALWAYS && !o.flags[v.flag]) continue;
        const int id = o.w.def_var(v.name, nc::NC_DOUBLE, dims_of(o.w.header(), v.dims));
        o.vid[k] = id;
        if (v.units) o.w.put_att_text(id, "Units", v.units);
        if (v.grid == 'b') {
            o.w.put_att(id, "N_grids", nc::NC_INT, &one, 1);
            o.w.put_att_text(id, "grid0", "tpts");
        } else if (v.grid) {
            o.w.put_att(id, "N_grids", nc::NC_INT, &three, 1);
            o.w.put_att_text(id, "grid0", "tpts");
            o.w.put_att_text(id, "grid1", v.grid == 'h' ? "ypts_edgH" : "ypts_avg");
            o.w.put_att_text(id, "grid2", v.grid == 'v' ? "xpts_edgV" : "xpts_avg");
        }
    }
}

// writes the context array of table entry k at time index t (1-based)
void write_entry(trt_outfile &o, size_t k, int t)
{
    const int id = o.vid[k];
    if (id < 0 || !o.table[k].arr) return;
    const nc::Var &v = o.w.header().vars[id];
    std::vector<double> a = pull(o.ctx, o.table[k].arr);
    std::vector<uint64_t> start(v.dims.size(), 0), count(v.dims.size());
    for (size_t d = 0; d < v.dims.size(); ++d) count[d] = o.w.header().dims[v.dims[d]].len;
    start[0] = (uint64_t)(t - 1);
    count[0] = 1;
    if (o.w.header().nelem(v) / o.nt != a.size()) throw Error(std::string("output: shape of ") + v.name);
    o.w.put(id, start.data(), count.data(), a.data());
}

void put_scalar_t(trt_outfile &o, const char *name, int t, double x)
{
    const int id = o.w.header().find_var(name);
    if (id < 0) return;
    const uint64_t s = (uint64_t)(t - 1), n = 1;
    o.w.put(id, &s, &n, &x);
}

void check_t(const trt_outfile &o, int t)
{
    if (t < 1 || t > o.nt) throw Error("output: time index " + std::to_string(t) + " outside 1.." + std::to_string(o.nt));
}

// the iteration counts and convergence norms (OUTFILE_VARDEFS :901-921), defined with
// dimensions as long as the longest iteration sequence written
void write_history(trt_outfile &o)
{
    int R = 1, M = 1, G = 1;
    for (auto &kv : o.hist)
        for (auto &r : kv.second) {
            R = std::max<int>(R, (int)kv.second.size());
            M = std::max<int>(M, (int)r.mgl.size());
            for (auto &m : r.mgl) G = std::max<int>(G, (int)m.gkits.size());
        }
    const int ng = o.ctx->c.ng;
    o.w.redef();
    const int dR = o.w.def_dim("RT_Its-dim", R), dM = o.w.def_dim("MGQD_Its-dim", M),
              dG = o.w.def_dim("GQD_Its-dim", G);
    const int dT = o.w.header().find_dim("N_t"), dg = o.w.header().find_dim("N_g");
    int rtc = -1, mgc = -1, gqc = -1, mgk = -1, gqk = -1;
    int rtn[4] = {-1, -1, -1, -1}, mgn[4] = {-1, -1, -1, -1};
    if (o.its_out) {
        rtc = o.w.def_var("RT_Its-count", nc::NC_INT, {dT});
        mgc = o.w.def_var("MGQD_Its-count", nc::NC_INT, {dT, dR});
        gqc = o.w.def_var("GQD_Its-count", nc::NC_INT, {dT, dR, dM});
        mgk = o.w.def_var("MGQD_MatVec_prds", nc::NC_INT, {dT, dR, dg});
        gqk = o.w.def_var("GQD_MatVec_prds", nc::NC_INT, {dT, dR, dM, dG});
    }
    if (o.conv_out) {
        // norms T, E then spectral radii T, E, as the reference orders them
        rtn[0] = o.w.def_var("RT_Tnorm", nc::NC_FLOAT, {dT, dR});
        rtn[1] = o.w.def_var("RT_Enorm", nc::NC_FLOAT, {dT, dR});
        mgn[0] = o.w.def_var("MGQD_Tnorm", nc::NC_FLOAT, {dT, dR, dM});
        mgn[1] = o.w.def_var("MGQD_Enorm", nc::NC_FLOAT, {dT, dR, dM});
        rtn[2] = o.w.def_var("RT_Trho", nc::NC_FLOAT, {dT, dR});
        rtn[3] = o.w.def_var("RT_Erho", nc::NC_FLOAT, {dT, dR});
        mgn[2] = o.w.def_var("MGQD_Trho", nc::NC_FLOAT, {dT, dR, dM});
        mgn[3] = o.w.def_var("MGQD_Erho", nc::NC_FLOAT, {dT, dR, dM});
    }
    o.w.enddef(o.fill, 0);
    for (auto &kv : o.hist) {
        const uint64_t t = (uint64_t)(kv.first - 1);
        const auto &rts = kv.second;
        if (o.its_out) {
            const int32_t n = (int32_t)rts.size();
            const uint64_t s1[1] = {t}, c1[1] = {1};
            o.w.put(rtc, s1, c1, &n);
        }
        for (size_t r = 0; r < rts.size(); ++r) {
            const auto &rl = rts[r];
            const uint64_t s2[2] = {t, r}, c2[2] = {1, 1};
            if (o.its_out) {
                const int32_t m = rl.mg;
                o.w.put(mgc, s2, c2, &m);
                if ((int)rl.kits.size() == ng) {
                    std::vector<int32_t> k(rl.kits.begin(), rl.kits.end());
                    const uint64_t s3[3] = {t, r, 0}, c3[3] = {1, 1, (uint64_t)ng};
                    o.w.put(mgk, s3, c3, k.data());
                }
            }
            if (o.conv_out) {
                const float v[4] = {(float)rl.tnorm, (float)rl.enorm, (float)rl.trho, (float)rl.erho};
                for (int q = 0; q < 4; ++q) o.w.put(rtn[q], s2, c2, &v[q]);
            }
            for (size_t m = 0; m < rl.mgl.size(); ++m) {
                const auto &ml = rl.mgl[m];
                const uint64_t s3[3] = {t, r, m}, c3[3] = {1, 1, 1};
                // the Newton counts exist only on the grey path (ALGORITHMS.f08:759-764)
                if (o.its_out && o.use_grey) {
                    const int32_t e = ml.newt;
                    o.w.put(gqc, s3, c3, &e);
                    if (!ml.gkits.empty()) {
                        std::vector<int32_t> k(ml.gkits.begin(), ml.gkits.end());
                        const uint64_t s4[4] = {t, r, m, 0}, c4[4] = {1, 1, 1, (uint64_t)k.size()};
                        o.w.put(gqk, s4, c4, k.data());
                    }
                }
                if (o.conv_out) {
                    const float v[4] = {(float)ml.tnorm, (float)ml.enorm, (float)ml.trho, (float)ml.erho};
                    for (int q = 0; q < 4; ++q) o.w.put(mgn[q], s3, c3, &v[q]);
                }
            }
        }
    }
}

}  // namespace

extern "C" {

int trt_outfile_create(trt_ctx *ctx, const char *path, const trt_outopts *op, trt_outfile **out)
{
    if (!out) return -1;
    *out = nullptr;
    if (!ctx || !path || !op) {
        trt::set_thread_error("trt_outfile_create: null argument");
        return -1;
    }
    trt_outfile *f = new trt_outfile;
    f->ctx = ctx;
    try {
        const trt::Ctx &c = ctx->c;
        const trt_problem &p = c.p;
        const trt_outopts &o = *op;
        if (o.kind != 0 && o.kind != 1) throw Error("output: kind must be 0 (run file) or 1 (Tgen_QDf file)");
        f->kind = o.kind;
        f->fill = !o.nofill;
        f->use_grey = p.no_grey ? 0 : 1;
        f->its_out = o.kind == 0 && o.its_out;
        f->conv_out = o.kind == 0 && o.conv_out;
        const int fl[F_ALWAYS + 1] = {o.I_out, o.HO_Eg_out, o.HO_Fg_out, o.HO_E_out, o.HO_F_out, o.Eg_out, o.Fg_out,
                                      o.MGQD_E_out, o.MGQD_F_out, o.QDfg_out, o.E_out, o.F_out, o.D_out,
                                      o.old_parms_out, o.kap_out, o.Src_out, 1};
        memcpy(f->flags, fl, sizeof fl);
        f->dx = pull(ctx, "Delx");
        f->dy = pull(ctx, "Dely");
        double xsum = 0, ysum = 0;
        for (double v : f->dx) xsum += v;
        for (double v : f->dy) ysum += v;
        const double xlen = o.xlen > 0 ? o.xlen : xsum, ylen = o.ylen > 0 ? o.ylen : ysum;
        f->ylen = ylen;
        // N_t and tpts as INPUT sets them (MainSrc/INPUTS.f08:120-151)
        int nt = o.n_t;
        if (nt <= 0) nt = (int)std::lround((o.tlen - o.start_time) / p.dt) + (o.init_out ? 1 : 0);
        if (nt < 1) throw Error("output: no time points (tlen, start_time and dt give N_t < 1)");
        f->nt = nt;
        std::vector<double> tpts(nt);
        if (o.tpts) {
            for (int i = 0; i < nt; ++i) tpts[i] = o.tpts[i];
        } else {
            tpts[0] = o.init_out ? o.start_time : o.start_time + p.dt;
            for (int i = 1; i < nt; ++i) tpts[i] = tpts[i - 1] + p.dt;
        }
        const int nx = c.nx, ny = c.ny;
        std::vector<double> xpa(nx), xpe(nx + 1), ypa(ny), ype(ny + 1);
        xpe[0] = 0.0;
        for (int i = 0; i < nx; ++i) xpe[i + 1] = xpe[i] + f->dx[i];
        xpa[0] = f->dx[0] / 2.0;
        for (int i = 1; i < nx; ++i) xpa[i] = xpa[i - 1] + (f->dx[i] + f->dx[i - 1]) / 2.0;
        ype[0] = 0.0;
        for (int j = 0; j < ny; ++j) ype[j + 1] = ype[j] + f->dy[j];
        ypa[0] = f->dy[0] / 2.0;
        for (int j = 1; j < ny; ++j) ypa[j] = ypa[j - 1] + (f->dy[j] + f->dy[j - 1]) / 2.0;

        if (o.format != 0 && o.format != 5) throw Error("output: format must be 0 or 5");
        nc::Writer &w = f->w;
        w.set_format(o.format);
        w.create(path);
        // dimensions (OUTPUTS.f08:51-66 / :432-442)
        const int dx_ = w.def_dim("N_x", nx), dy_ = w.def_dim("N_y", ny);
        w.def_dim("N_m", c.nm);
        w.def_dim("N_g", c.ng);
        const int dt_ = w.def_dim("N_t", nt), dV = w.def_dim("N_edgV", nx + 1), dH = w.def_dim("N_edgH", ny + 1);
        w.def_dim("N_xc", 2 * nx);
        w.def_dim("N_yc", 2 * ny);
        w.def_dim("Quadrants", 4);
        if (o.kind == 0) {
            // RT_Its-dim, MGQD_Its-dim and GQD_Its-dim (unlimited in the reference) are
            // defined at close, as long as the longest iteration sequence
            w.def_dim("Norm_Types", 5);
            w.def_dim("MGQD_ResTypes", 5);
        }
        w.def_dim("Boundaries", 4);
        // constants (OUTPUTS.f08:73-88)
        const double cval = p.c, hval = trt::H_PLANCK, pival = 4.0 * std::atan(1.0), ergval = trt::ERG;
        const int ic = w.def_var("c--lightspeed", nc::NC_DOUBLE, {});
        w.put_att_text(ic, "Units", "cm/sh");
        const int ih = w.def_var("h--Boltzmann_const", nc::NC_DOUBLE, {});
        w.put_att_text(ih, "Units", "erg*sh");
        const int icv = w.def_var("cv--specific_heat", nc::NC_DOUBLE, {});
        w.put_att_text(icv, "Units", "erg/(ev**4 cm**2 sh)");
        const int ie = w.def_var("ev<->erg_conversion", nc::NC_DOUBLE, {});
        w.put_att_text(ie, "Units", "eV/erg");
        const int ipi = w.def_var("pi", nc::NC_DOUBLE, {});
        const int icu = w.def_var("Computational_Unit", nc::NC_DOUBLE, {});
        auto ia = [&](const char *n, int v) { w.put_att(-1, n, nc::NC_INT, &v, 1); };
        auto da = [&](const char *n, double v) { w.put_att(-1, n, nc::NC_DOUBLE, &v, 1); };
        const int bct[4] = {p.bc_type[0], p.bc_type[1], p.bc_type[2], p.bc_type[3]};
        if (o.kind == 0) {
            // problem parameters as global attributes (OUTPUTS.f08:95-270)
            const std::string rt = text(o.run_type);
            w.put_att_text(-1, "run_type", rt);
            if (rt == "mg_pod" || rt == "mg_dmd") {
                const bool pod = rt == "mg_pod";
                w.put_att_text(-1, pod ? "POD_Type" : "DMD_Type", text(pod ? o.pod_type : o.dmd_type));
                if (pod) ia("Direc_Diff", o.direc_diff);
                if (o.n_dsets < 1) throw Error("output: ROM runs need n_dsets >= 1");
                w.def_dim("N_dsets", o.n_dsets);
                if (o.dset_times) w.put_att(-1, "dset_times", nc::NC_DOUBLE, o.dset_times, o.n_dsets);
                for (int i = 0; i < o.n_dsets; ++i)
                    w.put_att_text(-1, (pod ? "POD_dset" : "DMD_dset") + std::to_string(i + 1),
                                   text(o.dsets ? o.dsets[i] : nullptr));
            }
            ia("use_grey", f->use_grey);
            ia("maxit_RTE", p.maxit_rte);
            ia("maxit_MLOQD", p.maxit_mloqd);
            ia("maxit_GLOQD", p.maxit_gloqd);
            w.put_att_text(-1, "kapE_dT_flag", p.kapE_dT_flag ? "on" : "off");
            da("chi", p.chi);
            da("conv_ho", p.conv_ho);
            da("conv_lo", p.conv_lo);
            const double cg[2] = {p.conv_gr1, p.conv_gr2};
            w.put_att(-1, "conv_gr", nc::NC_DOUBLE, cg, 2);
            ia("conv_type", p.conv_type);
            w.put_att_text(-1, "quadrature", text(o.quadrature));
            ia("threads", p.threads);
            w.put_att_text(-1, "enrgy_strc", text(o.enrgy_strc));
            w.put_att(-1, "BC_type", nc::NC_INT, bct, 4);
            da("Theta", p.theta);
            if (p.use_line_search) {
                w.put_att_text(-1, "Use_Line_Search", "True");
                da("line_src", p.line_src);
            } else {
                w.put_att_text(-1, "Use_Line_Search", "False");
            }
            if (p.use_safety_search) {
                w.put_att_text(-1, "Use_Safety_Search", "True");
                da("E_Bound_Low", p.E_bound_low);
                da("T_Bound_Low", p.T_bound_low);
            } else {
                w.put_att_text(-1, "Use_Safety_Search", "False");
            }
            ia("I_out", o.I_out);
            ia("D_out", o.D_out);
            ia("fg_out", o.QDfg_out);
            ia("old_parms_out", o.old_parms_out);
            ia("HO_Eg_out", o.HO_Eg_out);
            ia("HO_E_out", o.HO_E_out);
            ia("HO_Fg_out", o.HO_Fg_out);
            ia("HO_F_out", o.HO_F_out);
            ia("Eg_out", o.Eg_out);
            ia("MGQD_E_out", o.MGQD_E_out);
            ia("Fg_out", o.Fg_out);
            ia("MGQD_F_out", o.MGQD_F_out);
            ia("E_out", o.E_out);
            ia("F_out", o.F_out);
            ia("its_out", o.its_out);
            ia("conv_out", o.conv_out);
            ia("kap_out", o.kap_out);
            ia("Src_out", o.Src_out);
        } else {
            // OUTFILE_INIT_Tgen's attributes (OUTPUTS.f08:470-503)
            w.put_att_text(-1, "T_run_type", text(o.run_type));
            w.put_att_text(-1, "quadrature", text(o.quadrature));
            ia("threads", p.threads);
            w.put_att_text(-1, "enrgy_strc", text(o.enrgy_strc));
            w.put_att(-1, "BC_type", nc::NC_INT, bct, 4);
            ia("I_out", o.I_out);
            ia("fg_out", o.QDfg_out);
            ia("HO_Eg_out", o.HO_Eg_out);
            ia("HO_E_out", o.HO_E_out);
            ia("HO_Fg_out", o.HO_Fg_out);
            ia("HO_F_out", o.HO_F_out);
        }
        // grid and problem variables (OUTPUTS.f08:272-307)
        auto scal = [&](const char *n, const char *u) {
            const int id = w.def_var(n, nc::NC_DOUBLE, {});
            w.put_att_text(id, "Units", u);
            return id;
        };
        const int ixl = scal("xlen", "cm"), iyl = scal("ylen", "cm");
        const int idx = w.def_var("Delx", nc::NC_DOUBLE, {dx_});
        w.put_att_text(idx, "Units", "cm");
        const int idy = w.def_var("Dely", nc::NC_DOUBLE, {dy_});
        w.put_att_text(idy, "Units", "cm");
        const int itl = scal("tlen", "sh"), idt = scal("Delt", "sh");
        const int ib0 = scal("bcT_left", "ev"), ib1 = scal("bcT_bottom", "ev"), ib2 = scal("bcT_right", "ev"),
                  ib3 = scal("bcT_top", "ev"), iti = scal("Tini", "ev");
        const double bx[2] = {0.0, xlen}, by[2] = {0.0, ylen}, bt[2] = {tpts[0], o.tlen};
        const int ixa = w.def_var("xpts_avg", nc::NC_DOUBLE, {dx_});
        w.put_att(ixa, "bnds", nc::NC_DOUBLE, bx, 2);
        const int ixe = w.def_var("xpts_edgV", nc::NC_DOUBLE, {dV});
        w.put_att(ixe, "bnds", nc::NC_DOUBLE, bx, 2);
        const int iya = w.def_var("ypts_avg", nc::NC_DOUBLE, {dy_});
        w.put_att(iya, "bnds", nc::NC_DOUBLE, by, 2);
        const int iye = w.def_var("ypts_edgH", nc::NC_DOUBLE, {dH});
        w.put_att(iye, "bnds", nc::NC_DOUBLE, by, 2);
        const int itp = w.def_var("tpts", nc::NC_DOUBLE, {dt_});
        w.put_att(itp, "bnds", nc::NC_DOUBLE, bt, 2);
        f->table = o.kind == 0 ? kRunVars : kTgenVars;
        f->ntable = o.kind == 0 ? sizeof(kRunVars) / sizeof(kRunVars[0]) : sizeof(kTgenVars) / sizeof(kTgenVars[0]);
        def_table(*f);
        // room in front of the data for the iteration dimensions and variables added at close
        w.enddef(f->fill, 16384);
        auto put0 = [&](int id, double v) { w.put(id, nullptr, nullptr, &v); };
        auto put1 = [&](int id, const std::vector<double> &v) {
            const uint64_t s = 0, n = v.size();
            w.put(id, &s, &n, v.data());
        };
        put0(ic, cval); put0(ih, hval); put0(ipi, pival); put0(ie, ergval); put0(icu, p.comp_unit); put0(icv, p.cv);
        put0(ixl, xlen); put0(iyl, ylen);
        put1(idx, f->dx); put1(idy, f->dy);
        put0(itl, o.tlen); put0(idt, p.dt);
        put0(ib0, p.bcT[0]); put0(ib1, p.bcT[1]); put0(ib2, p.bcT[2]); put0(ib3, p.bcT[3]); put0(iti, p.tini);
        put1(ixa, xpa); put1(ixe, xpe); put1(iya, ypa); put1(iye, ype); put1(itp, tpts);
    } catch (const Error &e) {
        trt::set_thread_error(e.msg);
        delete f;
        return -2;
    } catch (const std::exception &e) {
        trt::set_thread_error(e.what());
        delete f;
        return -2;
    }
    *out = f;
    return 0;
}

int trt_timestep_outs(trt_outfile *f, int t, int after_step)
{
    return out_guarded(f, [&](trt_outfile &o) {
        if (o.kind != 0) throw Error("output: trt_timestep_outs needs a run file (kind 0)");
        check_t(o, t);
        trt::Ctx &c = o.ctx->c;
        if (after_step) {
            // ALGORITHMS.f08:859-869: wave speeds of E and Temp, then the iteration history
            const std::vector<double> Ea = pull(o.ctx, "E_avg"), Ev = pull(o.ctx, "E_edgV"),
                                      Eo = pull(o.ctx, "E_avg_old"), T = pull(o.ctx, "Temp"),
                                      To = pull(o.ctx, "Temp_Old");
            put_scalar_t(o, "E_xWave_Speed", t,
                         xwave_speed_edg(Ea.data(), Ev.data(), Eo.data(), o.dy.data(), o.dx.data(), c.p.dt, o.ylen,
                                         c.ny, c.nx));
            put_scalar_t(o, "Temp_xWave_Speed", t,
                         xwave_speed_cnt(T.data(), To.data(), o.dy.data(), o.dx.data(), c.p.dt, o.ylen, c.ny, c.nx));
            if (o.its_out || o.conv_out) o.hist[t] = c.itlog;
        }
        for (size_t k = 0; k < o.ntable; ++k)
            if (!o.table[k].grey || o.use_grey) write_entry(o, k, t);
    });
}

int trt_tgen_outs(trt_outfile *f, int t)
{
    return out_guarded(f, [&](trt_outfile &o) {
        if (o.kind != 1) throw Error("output: trt_tgen_outs needs a Tgen_QDf file (kind 1)");
        check_t(o, t);
        for (size_t k = 0; k < o.ntable; ++k) write_entry(o, k, t);
    });
}

int trt_outfile_put(trt_outfile *f, const char *var, int ndims, const size_t *start, const size_t *count,
                    const double *vals)
{
    return out_guarded(f, [&](trt_outfile &o) {
        if (!var || !vals) throw Error("output: null argument");
        const nc::Header &h = o.w.header();
        const int id = h.find_var(var);
        if (id < 0) throw Error(std::string("output: no variable ") + var);
        const nc::Var &v = h.vars[id];
        if (ndims != (int)v.dims.size()) throw Error(std::string("output: rank of ") + var);
        // Fortran order and 1-based starts, as nf90_put_var takes them
        std::vector<uint64_t> s(ndims), n(ndims);
        uint64_t total = 1;
        for (int d = 0; d < ndims; ++d) {
            if (start[d] < 1) throw Error("output: starts are 1-based");
            s[ndims - 1 - d] = start[d] - 1;
            n[ndims - 1 - d] = count[d];
            total *= count[d];
        }
        const size_t es = nc::type_size(v.type);
        std::vector<unsigned char> buf(total * es);
        for (uint64_t i = 0; i < total; ++i) {
            if (v.type == nc::NC_DOUBLE) memcpy(&buf[i * 8], &vals[i], 8);
            else if (v.type == nc::NC_FLOAT) { float x = (float)vals[i]; memcpy(&buf[i * 4], &x, 4); }
            else if (v.type == nc::NC_INT) { int32_t x = (int32_t)std::lround(vals[i]); memcpy(&buf[i * 4], &x, 4); }
            else throw Error(std::string("output: ") + var + " has a type this call does not write");
        }
        o.w.put(id, s.data(), n.data(), buf.data());
    });
}

int trt_outfile_close(trt_outfile *f)
{
    if (!f) return -1;
    int rc = out_guarded(f, [&](trt_outfile &o) {
        if (o.kind == 0) write_history(o);
        o.w.close();
    });
    if (rc != 0) trt::set_thread_error(f->err);
    delete f;
    return rc;
}

const char *trt_outfile_last_error(const trt_outfile *f)
{
    return f ? f->err.c_str() : trt_last_error(nullptr);
}

int trt_nc_open(const char *path, trt_ncread **out)
{
    if (!out || !path) return -1;
    *out = nullptr;
    trt_ncread *r = new trt_ncread;
    try {
        r->r.open(path);
    } catch (const Error &e) {
        trt::set_thread_error(e.msg);
        delete r;
        return -2;
    }
    *out = r;
    return 0;
}

int trt_nc_dim(trt_ncread *r, const char *name, size_t *len)
{
    if (!r || !name || !len) return -1;
    const int d = r->r.header().find_dim(name);
    if (d < 0) {
        r->err = std::string("no dimension ") + name;
        return -2;
    }
    *len = (size_t)r->r.header().dims[d].len;
    return 0;
}

int trt_nc_get(trt_ncread *r, const char *var, int ndims, const size_t *start, const size_t *count, double *out)
{
    if (!r || !var || !out) return -1;
    try {
        const nc::Header &h = r->r.header();
        const int id = h.find_var(var);
        if (id < 0) throw Error(std::string("no variable ") + var);
        if (ndims != (int)h.vars[id].dims.size()) throw Error(std::string("rank of ") + var);
        std::vector<uint64_t> s(ndims), n(ndims);
        for (int d = 0; d < ndims; ++d) {
            if (start[d] < 1) throw Error("starts are 1-based");
            s[ndims - 1 - d] = start[d] - 1;
            n[ndims - 1 - d] = count[d];
        }
        r->r.get(id, s.data(), n.data(), out);
    } catch (const Error &e) {
        r->err = e.msg;
        return -2;
    }
    return 0;
}

int trt_nc_att(trt_ncread *r, const char *var, const char *name, double *vals, size_t cap, char *text, size_t tcap,
               size_t *n)
{
    if (!r || !name || !n) return -1;
    const nc::Header &h = r->r.header();
    const std::vector<nc::Att> *list = &h.gatts;
    if (var) {
        const int id = h.find_var(var);
        if (id < 0) {
            r->err = std::string("no variable ") + var;
            return -2;
        }
        list = &h.vars[id].atts;
    }
    for (const nc::Att &a : *list) {
        if (a.name != name) continue;
        *n = a.n;
        if (a.type == nc::NC_CHAR) {
            if (text && tcap) {
                const size_t k = std::min(a.n, tcap - 1);
                memcpy(text, a.raw.data(), k);
                text[k] = '\0';
            }
            return 1;
        }
        for (size_t i = 0; i < a.n && vals && i < cap; ++i) {
            const unsigned char *p = a.raw.data() + i * nc::type_size(a.type);
            double x;
            switch (a.type) {
            case nc::NC_BYTE: x = (double)*(const signed char *)p; break;
            case nc::NC_SHORT: { int16_t y; memcpy(&y, p, 2); x = y; break; }
            case nc::NC_INT: { int32_t y; memcpy(&y, p, 4); x = y; break; }
            case nc::NC_FLOAT: { float y; memcpy(&y, p, 4); x = y; break; }
            default: memcpy(&x, p, 8);
            }
            vals[i] = x;
        }
        return 0;
    }
    r->err = std::string("no attribute ") + name;
    return -2;
}

const char *trt_nc_last_error(const trt_ncread *r) { return r ? r->err.c_str() : trt_last_error(nullptr); }

int trt_nc_close(trt_ncread *r)
{
    if (!r) return -1;
    delete r;
    return 0;
}

}  // extern "C"
