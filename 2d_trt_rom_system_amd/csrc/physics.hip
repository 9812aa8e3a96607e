// K3 (material update) and K2 (moments -> Eddington/boundary QD factors).
//
// MATERIAL_UPDATE  MainSrc/UPDATES.f08:110-156
// kapB_calc / Bg_planck_calc / sigint  MainSrc/TEMP_FUNCTIONS.f08:11-141
// COLLAPSE_INTENSITIES  MainSrc/TRANSPORT_SOLVES.f08:381-511 (final reduction)
// fg_Calc :619-678, Cg_Calc :367-488, MGQD_In_Calc :493-614 of MainSrc/MLOQD_SOLVES.f08
// GQD_In_Calc  MainSrc/GLOQD_SOLVES.f08:263-290
#include "physics_device.h"

namespace trt {

struct MatArgs {
    int nx, ny, ng, model;
    double cu, inv_cdt;
    const double *T, *nu, *kmult;
    const int *mat;
    double *Bg, *KapB, *KapE, *KapR, *MGQD_Src, *RT_Src;
    double2 *sc;
    float2 *scf;   // FP32 copy for the FP32 sweep, or null
};

__global__ void material_kernel(MatArgs a)
{
    const size_t nc = (size_t)a.nx * a.ny;
    const size_t n = nc * a.ng;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x) {
        const size_t c = idx % nc;
        const int g = (int)(idx / nc);
        const double T = a.T[c];
        const double nl = a.nu[g], nh = a.nu[g + 1];
        const double bg = d_bg_planck(T, nl, nh, a.cu);
        const double kap = (a.model == 0 ? d_kapB(T, nl, nh) : d_kap_mid(T, nl, nh)) * a.kmult[a.mat[c]];
        a.Bg[idx] = bg;
        a.KapB[idx] = kap;
        a.KapE[idx] = kap;
        a.KapR[idx] = kap;
        a.MGQD_Src[idx] = 4.0 * PI * kap * bg;
        a.RT_Src[idx] = kap * bg;
        a.sc[c * a.ng + g] = make_double2(kap + a.inv_cdt, kap * bg);
        if (a.scf) a.scf[c * a.ng + g] = make_float2((float)(kap + a.inv_cdt), (float)(kap * bg * kIScale32));
    }
}

static int nblocks_for(size_t n, int threads = 256)
{
    size_t b = (n + threads - 1) / threads;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return (int)b;
}

void material_update(Ctx &c)
{
    MatArgs a{};
    a.nx = c.nx; a.ny = c.ny; a.ng = c.ng; a.model = c.p.opacity_model;
    a.cu = c.p.comp_unit;
    a.inv_cdt = 1.0 / (c.p.c * c.p.dt);
    a.T = c.arr("Temp"); a.nu = c.nu; a.kmult = c.kmult; a.mat = c.mat;
    a.Bg = c.arr("Bg"); a.KapB = c.arr("KapB"); a.KapE = c.arr("KapE"); a.KapR = c.arr("KapR");
    a.MGQD_Src = c.arr("MGQD_Src"); a.RT_Src = c.arr("RT_Src");
    a.sc = reinterpret_cast<double2 *>(c.sc);
    a.scf = reinterpret_cast<float2 *>(c.scf);
    const size_t n = (size_t)c.nx * c.ny * c.ng;
    hipLaunchKernelGGL(material_kernel, dim3(nblocks_for(n)), dim3(256), 0, c.stream, a);
    TRT_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// COLLAPSE_INTENSITIES: sum the 4 quadrant partials of every moment, and
// optionally fg_Calc's ratios (MLOQD_SOLVES.f08:646-674) in the same pass.
//
// The partials are stored by work item of the sweep (PartLayout: a cell's value v of
// quadrant q sits in q's item of the cell's tile, adjacent groups of a group block in
// adjacent doubles); the outputs are the reference's (location, group) arrays with the
// location fastest.  A block takes 64 locations x 4 groups of one location kind: it
// reads the partials with 4 adjacent lanes covering 4 adjacent groups, stages the
// results in LDS, and writes each output as 64 consecutive locations per group.
// ---------------------------------------------------------------------------
struct CollArgs {
    int nx, ny, ng;
    double cinv, c;
    const double *Part, *PVb, *PHb;         // quadrant partials (PartLayout) and incoming boundary faces
    PartLayout pl;
    double *Hxx, *Hyy, *Hxy, *Ea;           // cell averages
    double *Vxx, *Vxy, *Fx, *Ev;            // V faces
    double *Hyy2, *Hxy2, *Fy, *Eh;          // H faces
    double *fxx, *fyy, *fxy, *fvxx, *fvxy, *fhyy, *fhxy;   // fg_Calc outputs, or null
    int nbl[3];                             // location blocks of each kind
};

constexpr int CL_LOC = 64, CL_G = 4, CL_OUT = 7;

__global__ __launch_bounds__(256) void collapse_kernel(CollArgs a)
{
    __shared__ double st[CL_OUT][CL_G][CL_LOC];
    const int ngb = (a.ng + CL_G - 1) / CL_G;
    int b = blockIdx.x;
    int kind = 0;
    while (kind < 2 && b >= a.nbl[kind] * ngb) { b -= a.nbl[kind] * ngb; ++kind; }
    const int gb = b % ngb, lb = b / ngb;
    const size_t nloc = kind == 0 ? (size_t)a.nx * a.ny : kind == 1 ? (size_t)(a.nx + 1) * a.ny
                                                                    : (size_t)a.nx * (a.ny + 1);
    const int t = threadIdx.x;
    const bool fg = a.fxx != nullptr;
    {
        const int gl = t & (CL_G - 1), cl = t >> 2;
        const size_t loc = (size_t)lb * CL_LOC + cl;
        const int g = gb * CL_G + gl;
        if (loc < nloc && g < a.ng) {
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
            if (kind == 0) {
                const int i = (int)(loc % a.nx), j = (int)(loc / a.nx);
                for (int q = 0; q < 4; ++q) {
                    const double *p = a.Part + a.pl.at(i, j, q, g, 0);
                    const size_t PP = a.pl.PP;
                    s0 += p[0]; s1 += p[PP]; s2 += p[2 * PP]; s3 += p[3 * PP];
                }
            } else if (kind == 1) {
                const int f = (int)(loc % (a.nx + 1)), j = (int)(loc / (a.nx + 1));
                for (int q = 0; q < 4; ++q) {
                    s0 += part_v(a.pl, a.Part, a.PVb, f, j, q, g, 0);
                    s1 += part_v(a.pl, a.Part, a.PVb, f, j, q, g, 1);
                    s2 += part_v(a.pl, a.Part, a.PVb, f, j, q, g, 2);
                    s3 += part_v(a.pl, a.Part, a.PVb, f, j, q, g, 3);
                }
            } else {
                const int i = (int)(loc % a.nx), f = (int)(loc / a.nx);
                for (int q = 0; q < 4; ++q) {
                    s0 += part_h(a.pl, a.Part, a.PHb, i, f, q, g, 0);
                    s1 += part_h(a.pl, a.Part, a.PHb, i, f, q, g, 1);
                    s2 += part_h(a.pl, a.Part, a.PHb, i, f, q, g, 2);
                    s3 += part_h(a.pl, a.Part, a.PHb, i, f, q, g, 4);
                }
            }
            const double e = s3 * a.cinv;
            st[0][gl][cl] = s0; st[1][gl][cl] = s1; st[2][gl][cl] = s2; st[3][gl][cl] = e;
            if (fg) {
                const double ce = a.c * e;
                st[4][gl][cl] = s0 / ce;
                st[5][gl][cl] = s1 / ce;
                if (kind == 0) st[6][gl][cl] = s2 / ce;
            }
        }
    }
    __syncthreads();
    const int cl = t & (CL_LOC - 1), gl = t >> 6;
    const size_t loc = (size_t)lb * CL_LOC + cl;
    const int g = gb * CL_G + gl;
    if (loc >= nloc || g >= a.ng) return;
    const size_t o = loc + nloc * g;
    if (kind == 0) {
        a.Hxx[o] = st[0][gl][cl]; a.Hyy[o] = st[1][gl][cl]; a.Hxy[o] = st[2][gl][cl]; a.Ea[o] = st[3][gl][cl];
        if (fg) { a.fxx[o] = st[4][gl][cl]; a.fyy[o] = st[5][gl][cl]; a.fxy[o] = st[6][gl][cl]; }
    } else if (kind == 1) {
        a.Vxx[o] = st[0][gl][cl]; a.Vxy[o] = st[1][gl][cl]; a.Fx[o] = st[2][gl][cl]; a.Ev[o] = st[3][gl][cl];
        if (fg) { a.fvxx[o] = st[4][gl][cl]; a.fvxy[o] = st[5][gl][cl]; }
    } else {
        a.Hyy2[o] = st[0][gl][cl]; a.Hxy2[o] = st[1][gl][cl]; a.Fy[o] = st[2][gl][cl]; a.Eh[o] = st[3][gl][cl];
        if (fg) { a.fhyy[o] = st[4][gl][cl]; a.fhxy[o] = st[5][gl][cl]; }
    }
}

// grey sums over groups (TRANSPORT_SOLVES.f08:488-509 and COLLAPSE_MG_EF, MLOQD_SOLVES.f08:707-728)
struct GreySumArgs {
    int nx, ny, ng;
    const double *Ea, *Ev, *Eh, *Fx, *Fy;
    double *E_avg, *E_edgV, *E_edgH, *Fx_edgV, *Fy_edgH;
};

__global__ void grey_sum_kernel(GreySumArgs a)
{
    const size_t na = (size_t)a.nx * a.ny, nv = (size_t)(a.nx + 1) * a.ny, nh = (size_t)a.nx * (a.ny + 1);
    const size_t total = na + nv + nh;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        if (idx < na) {
            double t = 0.0;
            for (int g = 0; g < a.ng; ++g) t += a.Ea[idx + g * na];
            a.E_avg[idx] = t;
        } else if (idx < na + nv) {
            const size_t k = idx - na;
            double t = 0.0, f = 0.0;
            for (int g = 0; g < a.ng; ++g) { t += a.Ev[k + g * nv]; f += a.Fx[k + g * nv]; }
            a.E_edgV[k] = t; a.Fx_edgV[k] = f;
        } else {
            const size_t k = idx - na - nv;
            double t = 0.0, f = 0.0;
            for (int g = 0; g < a.ng; ++g) { t += a.Eh[k + g * nh]; f += a.Fy[k + g * nh]; }
            a.E_edgH[k] = t; a.Fy_edgH[k] = f;
        }
    }
}

void grey_sums(Ctx &c, const char *ea, const char *ev, const char *eh, const char *fx, const char *fy,
               const char *Ea, const char *Ev, const char *Eh, const char *Fx, const char *Fy)
{
    GreySumArgs a{};
    a.nx = c.nx; a.ny = c.ny; a.ng = c.ng;
    a.Ea = c.arr(ea); a.Ev = c.arr(ev); a.Eh = c.arr(eh); a.Fx = c.arr(fx); a.Fy = c.arr(fy);
    a.E_avg = c.arr(Ea); a.E_edgV = c.arr(Ev); a.E_edgH = c.arr(Eh); a.Fx_edgV = c.arr(Fx); a.Fy_edgH = c.arr(Fy);
    const size_t n = (size_t)c.nx * c.ny + (size_t)(c.nx + 1) * c.ny + (size_t)c.nx * (c.ny + 1);
    hipLaunchKernelGGL(grey_sum_kernel, dim3(nblocks_for(n)), dim3(256), 0, c.stream, a);
    TRT_CHECK(hipGetLastError());
}

static void collapse(Ctx &c, bool with_fg)
{
    CollArgs a{};
    a.nx = c.nx; a.ny = c.ny; a.ng = c.ng;
    a.cinv = 1.0 / C_CONST;     // TRANSPORT_SOLVES.f08:403,475-477 (hard-coded c)
    a.c = c.p.c;                // fg_Calc divides by c*E with the input c (MLOQD_SOLVES.f08:646-674)
    a.Part = c.Part; a.PVb = c.PVb; a.PHb = c.PHb; a.pl = c.pl;
    a.Hxx = c.arr("Hg_avg_xx"); a.Hyy = c.arr("Hg_avg_yy"); a.Hxy = c.arr("Hg_avg_xy"); a.Ea = c.arr("HO_Eg_avg");
    a.Vxx = c.arr("Hg_edgV_xx"); a.Vxy = c.arr("Hg_edgV_xy"); a.Fx = c.arr("HO_Fxg_edgV"); a.Ev = c.arr("HO_Eg_edgV");
    a.Hyy2 = c.arr("Hg_edgH_yy"); a.Hxy2 = c.arr("Hg_edgH_xy"); a.Fy = c.arr("HO_Fyg_edgH"); a.Eh = c.arr("HO_Eg_edgH");
    if (with_fg) {
        a.fxx = c.arr("fg_avg_xx"); a.fyy = c.arr("fg_avg_yy"); a.fxy = c.arr("fg_avg_xy");
        a.fvxx = c.arr("fg_edgV_xx"); a.fvxy = c.arr("fg_edgV_xy");
        a.fhyy = c.arr("fg_edgH_yy"); a.fhxy = c.arr("fg_edgH_xy");
    }
    const size_t nl[3] = {(size_t)c.nx * c.ny, (size_t)(c.nx + 1) * c.ny, (size_t)c.nx * (c.ny + 1)};
    const int ngb = (c.ng + CL_G - 1) / CL_G;
    size_t nblocks = 0;
    for (int k = 0; k < 3; ++k) {
        a.nbl[k] = (int)((nl[k] + CL_LOC - 1) / CL_LOC);
        nblocks += (size_t)a.nbl[k] * ngb;
    }
    if (nblocks > 0x7fffffff) throw Error("grid too large for the collapse kernel");
    hipLaunchKernelGGL(collapse_kernel, dim3((unsigned)nblocks), dim3(256), 0, c.stream, a);
    TRT_CHECK(hipGetLastError());
    grey_sums(c, "HO_Eg_avg", "HO_Eg_edgV", "HO_Eg_edgH", "HO_Fxg_edgV", "HO_Fyg_edgH",
              "HO_E_avg", "HO_E_edgV", "HO_E_edgH", "HO_Fx_edgV", "HO_Fy_edgH");
}

void collapse_intensities(Ctx &c) { collapse(c, false); }

void collapse_with_fg(Ctx &c) { collapse(c, true); }

// fg_Calc, MLOQD_SOLVES.f08:646-674
__global__ void ratio_kernel(const double *H, const double *E, double *f, double c, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        f[i] = H[i] / (c * E[i]);
}

void fg_calc(Ctx &c)
{
    const size_t na = (size_t)c.nx * c.ny * c.ng, nv = (size_t)(c.nx + 1) * c.ny * c.ng, nh = (size_t)c.nx * (c.ny + 1) * c.ng;
    const double cc = c.p.c;
    struct { const char *H, *E, *f; size_t n; } jobs[7] = {
        {"Hg_avg_xx", "HO_Eg_avg", "fg_avg_xx", na}, {"Hg_avg_yy", "HO_Eg_avg", "fg_avg_yy", na},
        {"Hg_avg_xy", "HO_Eg_avg", "fg_avg_xy", na}, {"Hg_edgV_xx", "HO_Eg_edgV", "fg_edgV_xx", nv},
        {"Hg_edgV_xy", "HO_Eg_edgV", "fg_edgV_xy", nv}, {"Hg_edgH_yy", "HO_Eg_edgH", "fg_edgH_yy", nh},
        {"Hg_edgH_xy", "HO_Eg_edgH", "fg_edgH_xy", nh}};
    for (auto &j : jobs)
        hipLaunchKernelGGL(ratio_kernel, dim3(nblocks_for(j.n)), dim3(256), 0, c.stream, c.arr(j.H), c.arr(j.E),
                           c.arr(j.f), cc, j.n);
    TRT_CHECK(hipGetLastError());
}

// Cg_Calc, MLOQD_SOLVES.f08:392-484: ratios of outgoing-half-range moments at the
// boundary faces, read from the quadrant partials (eps already included).
struct CgArgs {
    int nx, ny, ng, bc[4];
    const double *Part, *PVb, *PHb;
    PartLayout pl;
    double *L, *B, *R, *T;
};

__global__ void cg_kernel(CgArgs a)
{
    const int nside = 2 * a.ny + 2 * a.nx;
    const int total = nside * a.ng;
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += gridDim.x * blockDim.x) {
        const int g = idx / nside;
        int r = idx % nside;
        if (r < 2 * a.ny) {
            const bool left = r < a.ny;
            const int j = left ? r : r - a.ny;
            const int f = left ? 0 : a.nx;
            // left: outgoing Omega_x < 0 (q2, q3); right: Omega_x > 0 (q1, q4)
            const int qa = left ? 1 : 0, qb = left ? 2 : 3;
            const double v = (part_v(a.pl, a.Part, a.PVb, f, j, qa, g, 2) + part_v(a.pl, a.Part, a.PVb, f, j, qb, g, 2)) /
                             (part_v(a.pl, a.Part, a.PVb, f, j, qa, g, 3) + part_v(a.pl, a.Part, a.PVb, f, j, qb, g, 3));
            if (left) a.L[j + (size_t)a.ny * g] = a.bc[0] == 0 ? v : 0.0;
            else a.R[j + (size_t)a.ny * g] = a.bc[2] == 0 ? v : 0.0;
        } else {
            r -= 2 * a.ny;
            const bool bot = r < a.nx;
            const int i = bot ? r : r - a.nx;
            const int f = bot ? 0 : a.ny;
            // bottom: outgoing Omega_y < 0 (q3, q4); top: Omega_y > 0 (q1, q2)
            const int qa = bot ? 2 : 0, qb = bot ? 3 : 1;
            const double v = (part_h(a.pl, a.Part, a.PHb, i, f, qa, g, 3) + part_h(a.pl, a.Part, a.PHb, i, f, qb, g, 3)) /
                             (part_h(a.pl, a.Part, a.PHb, i, f, qa, g, 4) + part_h(a.pl, a.Part, a.PHb, i, f, qb, g, 4));
            if (bot) a.B[i + (size_t)a.nx * g] = a.bc[1] == 0 ? v : 0.0;
            else a.T[i + (size_t)a.nx * g] = a.bc[3] == 0 ? v : 0.0;
        }
    }
}

void cg_calc(Ctx &c)
{
    CgArgs a{};
    a.nx = c.nx; a.ny = c.ny; a.ng = c.ng;
    for (int k = 0; k < 4; ++k) a.bc[k] = c.p.bc_type[k];
    a.Part = c.Part; a.PVb = c.PVb; a.PHb = c.PHb; a.pl = c.pl;
    a.L = c.arr("Cg_L"); a.B = c.arr("Cg_B"); a.R = c.arr("Cg_R"); a.T = c.arr("Cg_T");
    const int n = (2 * c.nx + 2 * c.ny) * c.ng;
    hipLaunchKernelGGL(cg_kernel, dim3(nblocks_for(n)), dim3(256), 0, c.stream, a);
    TRT_CHECK(hipGetLastError());
}

// MGQD_In_Calc, MLOQD_SOLVES.f08:517-610: incoming half-range sums of the boundary
// intensities (no eps).  With incoming-Planckian boundaries the boundary
// intensity of every incoming direction is bcin[side][g].
struct InArgs {
    int nx, ny, nm, ng, bc[4];
    double c;
    const double *ox, *oy, *w, *bcin;
    double *EL, *EB, *ER, *ET, *FL, *FB, *FR, *FT;
};

__global__ void mgqd_in_kernel(InArgs a)
{
    const int nside = 2 * a.ny + 2 * a.nx;
    const int total = nside * a.ng;
    const int nm = a.nm;
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += gridDim.x * blockDim.x) {
        const int g = idx / nside;
        int r = idx % nside;
        double E = 0.0, F = 0.0;
        if (r < a.ny) {                 // left: m in [1,N_m/4] and (3N_m/4, N_m]
            const double I = a.bcin[0 * a.ng + g];
            if (a.bc[0] == 0) {
                for (int m = 0; m < nm / 4; ++m) { F += a.ox[m] * a.w[m] * I; E += a.w[m] * I; }
                for (int m = 3 * nm / 4; m < nm; ++m) { F += a.ox[m] * a.w[m] * I; E += a.w[m] * I; }
            }
            a.FL[r + (size_t)a.ny * g] = F; a.EL[r + (size_t)a.ny * g] = E / a.c;
        } else if (r < 2 * a.ny) {      // right: m in (N_m/4, 3N_m/4]
            const int j = r - a.ny;
            const double I = a.bcin[2 * a.ng + g];
            if (a.bc[2] == 0)
                for (int m = nm / 4; m < 3 * nm / 4; ++m) { F += a.ox[m] * a.w[m] * I; E += a.w[m] * I; }
            a.FR[j + (size_t)a.ny * g] = F; a.ER[j + (size_t)a.ny * g] = E / a.c;
        } else if (r < 2 * a.ny + a.nx) {   // bottom: m in [1, N_m/2]
            const int i = r - 2 * a.ny;
            const double I = a.bcin[1 * a.ng + g];
            if (a.bc[1] == 0)
                for (int m = 0; m < nm / 2; ++m) { F += a.oy[m] * a.w[m] * I; E += a.w[m] * I; }
            a.FB[i + (size_t)a.nx * g] = F; a.EB[i + (size_t)a.nx * g] = E / a.c;
        } else {                        // top: m in (N_m/2, N_m]
            const int i = r - 2 * a.ny - a.nx;
            const double I = a.bcin[3 * a.ng + g];
            if (a.bc[3] == 0)
                for (int m = nm / 2; m < nm; ++m) { F += a.oy[m] * a.w[m] * I; E += a.w[m] * I; }
            a.FT[i + (size_t)a.nx * g] = F; a.ET[i + (size_t)a.nx * g] = E / a.c;
        }
    }
}

void mgqd_in_calc(Ctx &c)
{
    InArgs a{};
    a.nx = c.nx; a.ny = c.ny; a.nm = c.nm; a.ng = c.ng;
    for (int k = 0; k < 4; ++k) a.bc[k] = c.p.bc_type[k];
    a.c = c.p.c;
    a.ox = c.ox; a.oy = c.oy; a.w = c.w; a.bcin = c.bcin;
    a.EL = c.arr("Eg_in_L"); a.EB = c.arr("Eg_in_B"); a.ER = c.arr("Eg_in_R"); a.ET = c.arr("Eg_in_T");
    a.FL = c.arr("Fg_in_L"); a.FB = c.arr("Fg_in_B"); a.FR = c.arr("Fg_in_R"); a.FT = c.arr("Fg_in_T");
    const int n = (2 * c.nx + 2 * c.ny) * c.ng;
    hipLaunchKernelGGL(mgqd_in_kernel, dim3(nblocks_for(n)), dim3(256), 0, c.stream, a);
    TRT_CHECK(hipGetLastError());
}

// GQD_In_Calc, GLOQD_SOLVES.f08:276-288
__global__ void sum_groups_kernel(const double *src, double *dst, int n, int ng)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double t = 0.0;
        for (int g = 0; g < ng; ++g) t += src[i + (size_t)n * g];
        dst[i] = t;
    }
}

void gqd_in_calc(Ctx &c)
{
    const char *mg[8] = {"Eg_in_L", "Eg_in_R", "Fg_in_L", "Fg_in_R", "Eg_in_B", "Eg_in_T", "Fg_in_B", "Fg_in_T"};
    const char *gr[8] = {"E_in_L", "E_in_R", "F_in_L", "F_in_R", "E_in_B", "E_in_T", "F_in_B", "F_in_T"};
    for (int k = 0; k < 8; ++k) {
        const int n = k < 4 ? c.ny : c.nx;
        hipLaunchKernelGGL(sum_groups_kernel, dim3(nblocks_for(n)), dim3(256), 0, c.stream, c.arr(mg[k]), c.arr(gr[k]),
                           n, c.ng);
    }
    TRT_CHECK(hipGetLastError());
}

void collapse_mg_ef(Ctx &c)
{
    grey_sums(c, "Eg_avg", "Eg_edgV", "Eg_edgH", "Fxg_edgV", "Fyg_edgH",
              "MGQD_E_avg", "MGQD_E_edgV", "MGQD_E_edgH", "MGQD_Fx_edgV", "MGQD_Fy_edgH");
}

}  // namespace trt
