// K1: simple-corner-balance discrete-ordinates sweep with fused angular moments.
//
// Replaces TRANSPORT_SCB (MainSrc/TRANSPORT_SOLVES.f08:14-327) and the angular
// sums of COLLAPSE_INTENSITIES (:381-511).
//
// Work decomposition (gfx950):
//  * Each quadrant is swept in its own reflected frame, where every direction
//    points to +x,+y.  SCB is reflection-symmetric: mirroring x maps the 4x4
//    system of Omega_x onto the one of -Omega_x with corners 1<->2, 3<->4
//    (SCB_BGEN, :332-371), so one kernel body serves the four quadrant loops
//    of :58-241 with a corner permutation at load/store time.
//  * A thread owns one (quadrant, group, direction) and sweeps the cells of a
//    TX x TY tile in upwind order, carrying the upwind corner-edge intensities
//    in registers.  Tiles on one tile-diagonal run concurrently (one launch per
//    diagonal); the faces between tiles go through small exchange buffers.
//  * A 256-thread block holds whole groups (gblk groups x K directions of the
//    quadrant), so the angular sums of COLLAPSE_INTENSITIES are finished per
//    block in LDS and written once per (cell or face, quadrant, group) in FP64:
//    4 quadrant partials instead of N_m/4 angle terms.
//  * Intensities are stored [cell][quadrant][group][direction][4 corners] so a
//    block's loads and stores of one cell are contiguous.
//  * The 4x4 corner solve uses the closed form of (gamma I + S)^-1 with S
//    skew-symmetric: S^2 = -(a^2+b^2) I + 2ab P with P^2 = I, hence
//    x = (gamma I - S)(alpha I + beta P) Q / ((gamma^2+(a-b)^2)(gamma^2+(a+b)^2)),
//    alpha = gamma^2+a^2+b^2, beta = 2ab.  cond(B) <= sqrt(2), so the FP32
//    variant solves in FP32 without loss beyond rounding.
#include "trt_internal.h"

#include <cstring>

namespace trt {

struct SweepArgs {
    int nx, ny, ng, K, gblk, nblk_g, TY, NTX, NTY, d, nt_d, tx_lo;
    double inv_cdt, eps;
    const double *dx, *dy, *ox, *oy, *w;
    const double2 *sc;      // [cell][g] (KapE + 1/(c dt), RT_Src)
    const double *bcin;     // [4][ng]
    const void *Iold;
    void *Inew;
    void *XV, *XH;
    double *Part, *PVb, *PHb;
    PartLayout pl;
};

template <typename T> struct Vec4;
template <> struct Vec4<double> {
    double v[4];
    __device__ static Vec4 load(const double *p) {
        const double2 a = *reinterpret_cast<const double2 *>(p);
        const double2 b = *reinterpret_cast<const double2 *>(p + 2);
        Vec4 r; r.v[0] = a.x; r.v[1] = a.y; r.v[2] = b.x; r.v[3] = b.y; return r;
    }
    __device__ void store(double *p) const {
        *reinterpret_cast<double2 *>(p) = make_double2(v[0], v[1]);
        *reinterpret_cast<double2 *>(p + 2) = make_double2(v[2], v[3]);
    }
};
template <> struct Vec4<float> {
    float v[4];
    __device__ static Vec4 load(const float *p) {
        const float4 a = *reinterpret_cast<const float4 *>(p);
        Vec4 r; r.v[0] = a.x; r.v[1] = a.y; r.v[2] = a.z; r.v[3] = a.w; return r;
    }
    __device__ void store(float *p) const {
        *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]);
    }
};

constexpr int SWEEP_THREADS = 256;

// Block-wide sum over the K directions of each group held by the block.
// vals[v] is this thread's contribution to moment v; out(v, g) receives the sum.
template <int NV, typename Out>
__device__ __forceinline__ void block_reduce(double *lds, const double (&vals)[NV], int K, int gblk, int g0,
                                             int ng, Out out)
{
    const int tid = threadIdx.x;
#pragma unroll
    for (int v = 0; v < NV; ++v) lds[v * SWEEP_THREADS + tid] = vals[v];
    __syncthreads();
    const int nout = gblk * NV;
    for (int o = tid; o < nout; o += SWEEP_THREADS) {
        const int gl = o / NV, v = o - gl * NV;
        const int g = g0 + gl;
        if (g < ng) {
            const double *src = lds + v * SWEEP_THREADS + gl * K;
            double s = 0.0;
            for (int k = 0; k < K; ++k) s += src[k];
            out(v, g, s);
        }
    }
    __syncthreads();
}

// One tile of one quadrant.  PX/PY: sign of Omega_x/Omega_y in this quadrant
// (q1 (+,+), q2 (-,+), q3 (-,-), q4 (+,-)).
template <typename TI, int TX, bool px, bool py>
__device__ __forceinline__ void sweep_tile(const SweepArgs &a, double *lds, int q, int gb, int txl, int tyl)
{
    const int tid = threadIdx.x;
    const int K = a.K, ng = a.ng, nx = a.nx, ny = a.ny;
    const int NL = ng * K;
    const int gl = tid / K, k = tid - (tid / K) * K;
    const int g0 = gb * a.gblk;
    const int g = g0 + gl;
    const bool act = (gl < a.gblk) && (g < ng);
    const int L = act ? g * K + k : 0;
    const int m = q * K + k;

    double sox = 0.0, soy = 0.0, wm = 0.0;
    if (act) { sox = a.ox[m]; soy = a.oy[m]; wm = a.w[m]; }
    const double axn = fabs(sox), ayn = fabs(soy);
    const double cw = wm, cxx = wm * sox * sox, cyy = wm * soy * soy, cxy = wm * sox * soy;
    const double cx = wm * sox, cy = wm * soy;
    // local corner (lb, rb, rt, lt) -> global corner (LB, RB, RT, LT)
    constexpr int pmask = (px ? 0 : 1) | (py ? 0 : 2);
    constexpr int perm[4] = {pmask == 0 ? 0 : pmask == 1 ? 1 : pmask == 2 ? 3 : 2,
                             pmask == 0 ? 1 : pmask == 1 ? 0 : pmask == 2 ? 2 : 3,
                             pmask == 0 ? 2 : pmask == 1 ? 3 : pmask == 2 ? 1 : 0,
                             pmask == 0 ? 3 : pmask == 1 ? 2 : pmask == 2 ? 0 : 1};

    const TI *Iold = static_cast<const TI *>(a.Iold);
    TI *Inew = static_cast<TI *>(a.Inew);
    TI *XV = static_cast<TI *>(a.XV);
    TI *XH = static_cast<TI *>(a.XH);
    const double eps = a.eps, isc = 1.0 / iscale<TI>();
    const TI inv_cdt = (TI)a.inv_cdt;

    const int i0 = txl * TX, iw = min(TX, nx - i0);
    const int j0 = tyl * a.TY, jh = min(a.TY, ny - j0);

    // incoming H corner-edge values of each column (local left, local right)
    TI hl[TX], hr[TX];
    {
        const TI bcy = act ? (TI)(a.bcin[(py ? 1 : 3) * ng + g] * iscale<TI>()) : (TI)0;
#pragma unroll
        for (int ii = 0; ii < TX; ++ii) {
            hl[ii] = bcy; hr[ii] = bcy;
            if (tyl > 0 && ii < iw && act) {
                const size_t base = ((size_t)(q * (a.NTY - 1) + tyl - 1) * (2 * nx) + 2 * (i0 + ii)) * NL + L;
                hl[ii] = XH[base];
                hr[ii] = XH[base + NL];
            }
        }
    }
    const TI bcx = act ? (TI)(a.bcin[(px ? 0 : 2) * ng + g] * iscale<TI>()) : (TI)0;

    for (int r = 0; r < jh; ++r) {
        const int jl = j0 + r;
        const int j = py ? jl : ny - 1 - jl;
        TI vb = bcx, vt = bcx;
        if (txl > 0 && act) {
            const size_t base = ((size_t)(q * (a.NTX - 1) + txl - 1) * (2 * ny) + 2 * jl) * NL + L;
            vb = XV[base];
            vt = XV[base + NL];
        }
        // prefetch the row
        Vec4<TI> io[TX];
        double2 scv[TX];
#pragma unroll
        for (int ii = 0; ii < TX; ++ii) {
            if (ii < iw && act) {
                const int il = i0 + ii;
                const int i = px ? il : nx - 1 - il;
                const size_t c = (size_t)i + (size_t)nx * j;
                io[ii] = Vec4<TI>::load(Iold + ((c * 4 + q) * NL + L) * 4);
                scv[ii] = a.sc[c * ng + g];
            } else {
                io[ii].v[0] = io[ii].v[1] = io[ii].v[2] = io[ii].v[3] = (TI)0;
                scv[ii] = make_double2(1.0, 0.0);
            }
        }
        const double dyc = a.dy[j];
#pragma unroll
        for (int ii = 0; ii < TX; ++ii) {
            if (ii >= iw) continue;
            const int il = i0 + ii;
            const int i = px ? il : nx - 1 - il;
            const size_t c = (size_t)i + (size_t)nx * j;
            const double dxc = a.dx[i];
            const TI A = (TI)(dxc * dyc);
            const TI A4 = A * (TI)0.25;
            const TI aa = (TI)(axn * dyc * 0.25), bb = (TI)(ayn * dxc * 0.25);
            const TI kt = (TI)scv[ii].x, src = (TI)(scv[ii].y * iscale<TI>());
            const TI gam = ((TI)(axn * dyc) + (TI)(ayn * dxc) + kt * A) * (TI)0.25;
            // local sources S = RT_Src + I_old / (c dt)  (TRANSPORT_SOLVES.f08:48)
            const TI s0 = src + io[ii].v[perm[0]] * inv_cdt;
            const TI s1 = src + io[ii].v[perm[1]] * inv_cdt;
            const TI s2 = src + io[ii].v[perm[2]] * inv_cdt;
            const TI s3 = src + io[ii].v[perm[3]] * inv_cdt;
            const TI Q0 = A4 * s0 + (TI)2 * aa * vb + (TI)2 * bb * hl[ii];
            const TI Q1 = A4 * s1 + (TI)2 * bb * hr[ii];
            const TI Q2 = A4 * s2;
            const TI Q3 = A4 * s3 + (TI)2 * aa * vt;
            const TI g2 = gam * gam;
            const TI al = g2 + aa * aa + bb * bb, be = (TI)2 * aa * bb;
            const TI y0 = al * Q0 + be * Q2, y1 = al * Q1 - be * Q3;
            const TI y2 = al * Q2 + be * Q0, y3 = al * Q3 - be * Q1;
            const TI dm = aa - bb, dp = aa + bb;
            const TI rD = (TI)1 / ((g2 + dm * dm) * (g2 + dp * dp));
            const TI x0 = (gam * y0 - aa * y1 - bb * y3) * rD;
            const TI x1 = (gam * y1 + aa * y0 - bb * y2) * rD;
            const TI x2 = (gam * y2 + bb * y1 + aa * y3) * rD;
            const TI x3 = (gam * y3 + bb * y0 - aa * y2) * rD;
            const double IVin = ((double)vb + (double)vt) * 0.5 * isc;
            const double IHin = ((double)hl[ii] + (double)hr[ii]) * 0.5 * isc;
            if (act) {
                Vec4<TI> o;
                o.v[perm[0]] = x0; o.v[perm[1]] = x1; o.v[perm[2]] = x2; o.v[perm[3]] = x3;
                o.store(Inew + ((c * 4 + q) * NL + L) * 4);
            }
            // upwinding to the outgoing corner edges (TRANSPORT_SOLVES.f08:86-89)
            vb = x1; vt = x2; hl[ii] = x3; hr[ii] = x2;
            // angular moments (COLLAPSE_INTENSITIES, :447-470), eps added in FP64
            const double Ia = ((double)x0 + (double)x1 + (double)x2 + (double)x3) * 0.25 * isc + eps;
            const double Iv = ((double)x1 + (double)x2) * 0.5 * isc + eps;
            const double Ih = ((double)x3 + (double)x2) * 0.5 * isc + eps;
            const double vals[NMOM] = {cxx * Ia, cyy * Ia, cxy * Ia, cw * Ia,
                                       cxx * Iv, cxy * Iv, cx * Iv, cw * Iv,
                                       cyy * Ih, cxy * Ih, cx * Ih, cy * Ih, cw * Ih};
            block_reduce<NMOM>(lds, vals, K, a.gblk, g0, ng, [&](int v, int gg, double s) {
                a.Part[a.pl.at(i, j, q, gg, v)] = s;
            });
            // incoming boundary faces contribute their own moments (their values are the
            // boundary intensities set by RT_BC_UPDATE, UPDATES.f08:12-105)
            const bool xin = (txl == 0 && ii == 0), yin = (tyl == 0 && r == 0);
            if (xin || yin) {
                const double Ivb = IVin + eps, Ihb = IHin + eps;
                const double bv[9] = {cxx * Ivb, cxy * Ivb, cx * Ivb, cw * Ivb,
                                      cyy * Ihb, cxy * Ihb, cx * Ihb, cy * Ihb, cw * Ihb};
                block_reduce<9>(lds, bv, K, a.gblk, g0, ng, [&](int v, int gg, double s) {
                    if (v < 4) { if (xin) a.PVb[(((size_t)q * ny + j) * ng + gg) * 4 + v] = s; }
                    else { if (yin) a.PHb[(((size_t)q * nx + i) * ng + gg) * 5 + (v - 4)] = s; }
                });
            }
        }
        if (txl < a.NTX - 1 && act) {
            const size_t base = ((size_t)(q * (a.NTX - 1) + txl) * (2 * ny) + 2 * jl) * NL + L;
            XV[base] = vb;
            XV[base + NL] = vt;
        }
    }
    if (tyl < a.NTY - 1 && act) {
#pragma unroll
        for (int ii = 0; ii < TX; ++ii) {
            if (ii < iw) {
                const size_t base = ((size_t)(q * (a.NTY - 1) + tyl) * (2 * nx) + 2 * (i0 + ii)) * NL + L;
                XH[base] = hl[ii];
                XH[base + NL] = hr[ii];
            }
        }
    }
}

template <typename TI, int TX>
__global__ __launch_bounds__(SWEEP_THREADS) void sweep_kernel(SweepArgs a)
{
    __shared__ double lds[NMOM * SWEEP_THREADS];
    int bid = blockIdx.x;
    const int gb = bid % a.nblk_g;
    bid /= a.nblk_g;
    const int t = bid % a.nt_d;
    const int q = bid / a.nt_d;
    const int txl = a.tx_lo + t, tyl = a.d - txl;
    switch (q) {
    case 0: sweep_tile<TI, TX, true, true>(a, lds, q, gb, txl, tyl); break;
    case 1: sweep_tile<TI, TX, false, true>(a, lds, q, gb, txl, tyl); break;
    case 2: sweep_tile<TI, TX, false, false>(a, lds, q, gb, txl, tyl); break;
    default: sweep_tile<TI, TX, true, false>(a, lds, q, gb, txl, tyl); break;
    }
}

static void geom_for(const trt_problem &p, SweepGeom &sg)
{
    sg.K = p.nm / 4;
    // as few group blocks as 256 threads allow, with the groups spread evenly over them
    // (32 groups at S8: 16 + 16 rather than 25 + 7), so that the blocks' dependency
    // chains carry equal work and no block is mostly padding
    const int gmax = std::min(SWEEP_THREADS / sg.K, p.ng);
    sg.nblk_g = (p.ng + gmax - 1) / gmax;
    sg.gblk = (p.ng + sg.nblk_g - 1) / sg.nblk_g;
    // default tiles: 4x4 for sweep_mode 1; for the tile engine 4x2 (FP32), 2x4 (FP64) and
    // 2x2 (FP32 storage with FP64 arithmetic), the fastest measured on the 256x256, S16,
    // 32-group problem
    const bool engine = p.sweep_mode != 1;
    // default tiles: FP32 4x4, FP64 2x4, FP32 storage with FP64 arithmetic 2x2 (the per-row
    // form of sweep_mode 1: 4x4); trt_autotune_tile times the others
    sg.TX = p.tile_x > 0 ? p.tile_x : (!engine ? 4 : p.precision == 1 ? 4 : 2);
    sg.TY = p.tile_y > 0 ? p.tile_y : (!engine ? 4 : p.precision == 2 ? 2 : 4);
    sg.NTX = (p.nx + sg.TX - 1) / sg.TX;
    sg.NTY = (p.ny + sg.TY - 1) / sg.TY;
}

static PartLayout part_layout(const trt_problem &p, const SweepGeom &sg)
{
    PartLayout L;
    L.nx = p.nx; L.ny = p.ny; L.ng = p.ng;
    L.TX = sg.TX; L.TY = sg.TY; L.NTX = sg.NTX; L.NTY = sg.NTY;
    L.gblk = sg.gblk; L.nblk_g = sg.nblk_g;
    L.PP = (sg.TX * sg.TY * sg.gblk + 15) / 16 * 16;   // runs of whole 128-byte lines
    return L;
}

size_t sweep_bytes(const trt_problem &p, SweepGeom *out)
{
    SweepGeom sg;
    geom_for(p, sg);
    if (out) *out = sg;
    const size_t nc = (size_t)p.nx * p.ny, NL = (size_t)p.ng * sg.K;
    const size_t ti = p.precision ? 4 : 8;
    size_t b = 2 * nc * 4 * NL * 4 * ti;                                   // I_crn, I_crn_old
    const size_t NR = (size_t)rec_lanes((int)NL);
    b += 4 * (size_t)std::max(sg.NTX - 1, 0) * 2 * sg.NTY * sg.TY * NR * ti;   // XV (whole tiles)
    b += 4 * (size_t)std::max(sg.NTY - 1, 0) * 2 * sg.NTX * sg.TX * NR * ti;   // XH
    bool refl = false;
    for (int k = 0; k < 4; ++k) refl = refl || p.bc_type[k] == 1;
    if (refl)                                                              // boundary slices
        b += 2 * 4 * ((size_t)sg.NTY * 2 * sg.TY + (size_t)sg.NTX * 2 * sg.TX) * NR * ti;
    b += (size_t)p.nm * 6 * 8;                                             // moment weight table
    b += nc * p.ng * 2 * 8;                                                // sc
    if (p.precision) b += nc * p.ng * 2 * 4;                               // FP32 copy of sc
    b += (part_layout(p, sg).size() + 4 * (size_t)p.ng * (p.ny * 4 + p.nx * 5)) * 8;   // partial moments
    return b;
}

void sweep_setup(Ctx &c)
{
    sweep_bytes(c.p, &c.sg);
    const SweepGeom &sg = c.sg;
    if (c.nm % 4) throw Error("N_m must be a multiple of 4 (quadrant blocks)");
    if (sg.K > SWEEP_THREADS) throw Error("more than 256 directions per quadrant is not supported");
    c.sweep_mode = c.p.sweep_mode == 0 ? 2 : c.p.sweep_mode;
    if (c.sweep_mode == 1 && sg.TX != 4 && sg.TX != 8 && sg.TX != 16)
        throw Error("tile_x must be 4, 8 or 16 for sweep_mode 1");
    if (c.sweep_mode >= 2 && !pk_supported(c))
        throw Error("sweep_mode 2, 3 and 4 take tiles 2x2, 2x4, 4x2 (FP64) and also 4x4, 8x2, 2x8 (FP32)");
    const size_t nc = (size_t)c.nx * c.ny, NL = (size_t)c.ng * sg.K;
    const size_t ti = c.precision ? 4 : 8;
    c.icrn_bytes = nc * 4 * NL * 4 * ti;
    c.Icrn[0] = c.dalloc(c.icrn_bytes);
    c.Icrn[1] = c.dalloc(c.icrn_bytes);
    // face exchange: sized for whole tiles (the tile engine's packed layout), which also
    // covers the per-row layout of sweep_mode 1
    const size_t NR = (size_t)rec_lanes((int)NL);      // also covers the per-row layout's NL lanes
    c.XV = c.dalloc(std::max<size_t>(1, 4 * (size_t)(sg.NTX - 1) * 2 * sg.NTY * sg.TY * NR * ti));
    c.XH = c.dalloc(std::max<size_t>(1, 4 * (size_t)(sg.NTY - 1) * 2 * sg.NTX * sg.TX * NR * ti));
    c.refl = 0;
    for (int k = 0; k < 4; ++k)
        if (c.p.bc_type[k] == 1) c.refl |= 1 << k;
    if (c.refl && c.sweep_mode == 1)
        throw Error("reflective boundaries (BC_Type 1) run on the tile engine (sweep_mode 0, 2, 3 or 4)");
    if (c.refl) {
        c.BV = c.dalloc(2 * 4 * (size_t)sg.NTY * NR * 2 * sg.TY * ti);
        c.BH = c.dalloc(2 * 4 * (size_t)sg.NTX * NR * 2 * sg.TX * ti);
    }
    c.sc = static_cast<double *>(c.dalloc(nc * c.ng * 2 * 8));
    if (c.precision) c.scf = static_cast<float *>(c.dalloc(nc * c.ng * 2 * 4));
    c.pl = part_layout(c.p, sg);
    c.Part = static_cast<double *>(c.dalloc(c.pl.size() * 8));
    c.PVb = static_cast<double *>(c.dalloc((size_t)4 * c.ny * c.ng * 4 * 8));
    c.PHb = static_cast<double *>(c.dalloc((size_t)4 * c.nx * c.ng * 5 * 8));
}

template <typename TI>
static void launch_sweep(Ctx &c, SweepArgs &a, int nblocks)
{
    switch (c.sg.TX) {
    case 4: hipLaunchKernelGGL((sweep_kernel<TI, 4>), dim3(nblocks), dim3(SWEEP_THREADS), 0, c.stream, a); break;
    case 8: hipLaunchKernelGGL((sweep_kernel<TI, 8>), dim3(nblocks), dim3(SWEEP_THREADS), 0, c.stream, a); break;
    default: hipLaunchKernelGGL((sweep_kernel<TI, 16>), dim3(nblocks), dim3(SWEEP_THREADS), 0, c.stream, a); break;
    }
}

void transport_sweep(Ctx &c, std::vector<hipEvent_t> *ev)
{
    const SweepGeom &sg = c.sg;
    if (c.sweep_mode >= 2) {
        const int out = 1 - c.iold;
        if (c.sweep_mode == 2 || c.sweep_mode == 4) pk_sweep(c, ev);
        else diag_sweep(c, ev);
        TRT_CHECK(hipGetLastError());
        c.icur = out;
        return;
    }
    // the new intensities go to the buffer that does not hold I_crn_old
    const int out = 1 - c.iold;
    SweepArgs a{};
    a.nx = c.nx; a.ny = c.ny; a.ng = c.ng; a.K = sg.K; a.gblk = sg.gblk; a.nblk_g = sg.nblk_g;
    a.TY = sg.TY; a.NTX = sg.NTX; a.NTY = sg.NTY;
    a.inv_cdt = 1.0 / (c.p.c * c.p.dt);
    a.eps = 1e-25 / c.p.comp_unit;
    a.dx = c.dx; a.dy = c.dy; a.ox = c.ox; a.oy = c.oy; a.w = c.w;
    a.sc = reinterpret_cast<const double2 *>(c.sc);
    a.bcin = c.bcin;
    a.Iold = c.Icrn[c.iold];
    a.Inew = c.Icrn[out];
    a.XV = c.XV; a.XH = c.XH;
    a.Part = c.Part; a.PVb = c.PVb; a.PHb = c.PHb; a.pl = c.pl;
    for (int d = 0; d <= sg.NTX + sg.NTY - 2; ++d) {
        a.d = d;
        a.tx_lo = std::max(0, d - (sg.NTY - 1));
        const int tx_hi = std::min(d, sg.NTX - 1);
        a.nt_d = tx_hi - a.tx_lo + 1;
        const int nblocks = 4 * a.nt_d * sg.nblk_g;
        if (ev) TRT_CHECK(hipEventRecord((*ev)[2 * d], c.stream));
        if (c.precision) launch_sweep<float>(c, a, nblocks);
        else launch_sweep<double>(c, a, nblocks);
        if (ev) TRT_CHECK(hipEventRecord((*ev)[2 * d + 1], c.stream));
    }
    TRT_CHECK(hipGetLastError());
    c.icur = out;
}

// ---------------------------------------------------------------------------
// layout conversion between the reference's I_crn(2N_x,2N_y,N_m,N_g) and the
// device layout [cell][q][g][k][corner]
// ---------------------------------------------------------------------------
template <typename TI>
__global__ void icrn_to_fortran(const TI *src, double *dst, int nx, int ny, int nm, int ng)
{
    const size_t n = (size_t)4 * nx * ny * nm * ng;
    const int K = nm / 4;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x) {
        size_t r = idx;
        const int ii = r % (2 * nx); r /= (2 * nx);
        const int jj = r % (2 * ny); r /= (2 * ny);
        const int m = r % nm; const int g = (int)(r / nm);
        const int i = ii / 2, j = jj / 2, right = ii & 1, top = jj & 1;
        const int corner = top ? (right ? 2 : 3) : (right ? 1 : 0);
        const int q = m / K, k = m % K;
        const size_t c = (size_t)i + (size_t)nx * j;
        dst[idx] = (double)src[((c * 4 + q) * ((size_t)ng * K) + (size_t)g * K + k) * 4 + corner] * (1.0 / iscale<TI>());
    }
}

template <typename TI>
__global__ void icrn_from_fortran(const double *src, TI *dst, int nx, int ny, int nm, int ng)
{
    const size_t n = (size_t)4 * nx * ny * nm * ng;
    const int K = nm / 4;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x) {
        size_t r = idx;
        const int ii = r % (2 * nx); r /= (2 * nx);
        const int jj = r % (2 * ny); r /= (2 * ny);
        const int m = r % nm; const int g = (int)(r / nm);
        const int i = ii / 2, j = jj / 2, right = ii & 1, top = jj & 1;
        const int corner = top ? (right ? 2 : 3) : (right ? 1 : 0);
        const int q = m / K, k = m % K;
        const size_t c = (size_t)i + (size_t)nx * j;
        dst[((c * 4 + q) * ((size_t)ng * K) + (size_t)g * K + k) * 4 + corner] = (TI)(src[idx] * iscale<TI>());
    }
}

void pull_icrn(Ctx &c, int which, double *host)
{
    const size_t n = (size_t)4 * c.nx * c.ny * c.nm * c.ng;
    double *tmp = static_cast<double *>(c.dalloc(n * 8));
    if (c.precision)
        hipLaunchKernelGGL(icrn_to_fortran<float>, dim3(1024), dim3(256), 0, c.stream,
                           static_cast<const float *>(c.Icrn[which]), tmp, c.nx, c.ny, c.nm, c.ng);
    else
        hipLaunchKernelGGL(icrn_to_fortran<double>, dim3(1024), dim3(256), 0, c.stream,
                           static_cast<const double *>(c.Icrn[which]), tmp, c.nx, c.ny, c.nm, c.ng);
    TRT_CHECK(hipMemcpyAsync(host, tmp, n * 8, hipMemcpyDeviceToHost, c.stream));
    TRT_CHECK(hipStreamSynchronize(c.stream));
    c.dfree_last();
}

void push_icrn(Ctx &c, int which, const double *host)
{
    const size_t n = (size_t)4 * c.nx * c.ny * c.nm * c.ng;
    double *tmp = static_cast<double *>(c.dalloc(n * 8));
    TRT_CHECK(hipMemcpyAsync(tmp, host, n * 8, hipMemcpyHostToDevice, c.stream));
    if (c.precision)
        hipLaunchKernelGGL(icrn_from_fortran<float>, dim3(1024), dim3(256), 0, c.stream, tmp,
                           static_cast<float *>(c.Icrn[which]), c.nx, c.ny, c.nm, c.ng);
    else
        hipLaunchKernelGGL(icrn_from_fortran<double>, dim3(1024), dim3(256), 0, c.stream, tmp,
                           static_cast<double *>(c.Icrn[which]), c.nx, c.ny, c.nm, c.ng);
    TRT_CHECK(hipStreamSynchronize(c.stream));
    c.dfree_last();
}


// ---------------------------------------------------------------------------
// Reference-shaped views derived from I_crn (pull only): I_avg(N_x,N_y,N_m,N_g),
// I_edgV(N_x+1,N_y,..), I_edgH(N_x,N_y+1,..), Ic_edgV(N_x+1,2N_y,..), Ic_edgH(2N_x,N_y+1,..).
// Interior and outgoing faces hold the upwind corner values (TRANSPORT_SOLVES.f08:86-93),
// incoming boundary faces the boundary intensity (RT_BC_UPDATE, UPDATES.f08:26-102): the
// Planckian of the side, or on a reflective side the inflow slice (after RT_BC_UPDATE, the
// mirror direction's outflow of the last sweep).
// ---------------------------------------------------------------------------
struct InflowSlices {
    const void *BV, *BH;
    int refl, TX, TY, NTX, NTY, NLr;
};

template <typename TI>
__global__ void derived_kernel(const TI *I, const double *bcin, InflowSlices bs, double *out, int kind, int nx, int ny,
                               int nm, int ng)
{
    const int K = nm / 4;
    const size_t NL = (size_t)ng * K;
    int d0, d1;
    switch (kind) {
    case 0: d0 = nx; d1 = ny; break;
    case 1: d0 = nx + 1; d1 = ny; break;
    case 2: d0 = nx; d1 = ny + 1; break;
    case 3: d0 = nx + 1; d1 = 2 * ny; break;
    default: d0 = 2 * nx; d1 = ny + 1; break;
    }
    const size_t n = (size_t)d0 * d1 * nm * ng;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x) {
        size_t r = idx;
        const int a = r % d0; r /= d0;
        const int b = r % d1; r /= d1;
        const int m = r % nm;
        const int g = (int)(r / nm);
        const int q = m / K, k = m % K;
        const bool px = (q == 0 || q == 3), py = (q < 2);
        auto rec = [&](int i, int j) { return I + (((size_t)i + (size_t)nx * j) * 4 + q) * NL * 4 + ((size_t)g * K + k) * 4; };
        double v = 0.0;
        if (kind == 0) {
            const TI *p = rec(a, b);
            v = ((double)p[0] + (double)p[1] + (double)p[2] + (double)p[3]) * 0.25 * (1.0 / iscale<TI>());
        } else if (kind == 1 || kind == 3) {          // vertical faces
            const int f = a;
            const int j = kind == 1 ? b : b / 2, top = kind == 3 ? (b & 1) : -1;
            if ((px && f == 0) || (!px && f == nx)) {
                const int side = px ? 0 : 2;
                if (bs.refl & (1 << side)) {
                    const int jf = py ? j : ny - 1 - j, tile = jf / bs.TY, r = jf - tile * bs.TY;
                    const TI *sl = static_cast<const TI *>(bs.BV) + (size_t)((px ? 0 : 1) * 4 + q) * bs.NTY * bs.NLr * 2 * bs.TY;
                    const int lane = g * K + k;
                    const double sc = 1.0 / iscale<TI>();
                    // frame-bottom and frame-top corners; the frame is flipped in y when !py
                    const double fb = (double)sl[rec_elem<TI>(tile, 2 * r, 2 * bs.TY, bs.NLr, lane)] * sc;
                    const double ft = (double)sl[rec_elem<TI>(tile, 2 * r + 1, 2 * bs.TY, bs.NLr, lane)] * sc;
                    const double lo = py ? fb : ft, hi = py ? ft : fb;
                    v = top < 0 ? (lo + hi) * 0.5 : (top ? hi : lo);
                } else {
                    v = bcin[side * ng + g];
                }
            } else {
                const TI *p = px ? rec(f - 1, j) : rec(f, j);
                const double sc = 1.0 / iscale<TI>();
                const double lo = (px ? (double)p[1] : (double)p[0]) * sc, hi = (px ? (double)p[2] : (double)p[3]) * sc;
                v = top < 0 ? (lo + hi) * 0.5 : (top ? hi : lo);
            }
        } else {                                      // horizontal faces
            const int f = b;
            const int i = kind == 2 ? a : a / 2, right = kind == 4 ? (a & 1) : -1;
            if ((py && f == 0) || (!py && f == ny)) {
                const int side = py ? 1 : 3;
                if (bs.refl & (1 << side)) {
                    const int xf = px ? i : nx - 1 - i, tile = xf / bs.TX, ii = xf - tile * bs.TX;
                    const TI *sl = static_cast<const TI *>(bs.BH) + (size_t)((py ? 0 : 1) * 4 + q) * bs.NTX * bs.NLr * 2 * bs.TX;
                    const int lane = g * K + k;
                    const double sc = 1.0 / iscale<TI>();
                    // frame-left and frame-right corners; the frame is flipped in x when !px
                    const double fl = (double)sl[rec_elem<TI>(tile, 2 * ii, 2 * bs.TX, bs.NLr, lane)] * sc;
                    const double fr = (double)sl[rec_elem<TI>(tile, 2 * ii + 1, 2 * bs.TX, bs.NLr, lane)] * sc;
                    const double lft = px ? fl : fr, rgt = px ? fr : fl;
                    v = right < 0 ? (lft + rgt) * 0.5 : (right ? rgt : lft);
                } else {
                    v = bcin[side * ng + g];
                }
            } else {
                const TI *p = py ? rec(i, f - 1) : rec(i, f);
                const double sc = 1.0 / iscale<TI>();
                const double lft = (py ? (double)p[3] : (double)p[0]) * sc, rgt = (py ? (double)p[2] : (double)p[1]) * sc;
                v = right < 0 ? (lft + rgt) * 0.5 : (right ? rgt : lft);
            }
        }
        out[idx] = v;
    }
}

size_t derived_size(const Ctx &c, const char *name)
{
    const size_t mg = (size_t)c.nm * c.ng;
    if (!strcmp(name, "I_avg")) return (size_t)c.nx * c.ny * mg;
    if (!strcmp(name, "I_edgV")) return (size_t)(c.nx + 1) * c.ny * mg;
    if (!strcmp(name, "I_edgH")) return (size_t)c.nx * (c.ny + 1) * mg;
    if (!strcmp(name, "Ic_edgV")) return (size_t)(c.nx + 1) * 2 * c.ny * mg;
    if (!strcmp(name, "Ic_edgH")) return (size_t)2 * c.nx * (c.ny + 1) * mg;
    return 0;
}

void pull_derived(Ctx &c, const char *name, double *host)
{
    int kind = -1;
    const char *names[5] = {"I_avg", "I_edgV", "I_edgH", "Ic_edgV", "Ic_edgH"};
    for (int k = 0; k < 5; ++k) if (!strcmp(name, names[k])) kind = k;
    if (kind < 0) throw Error(std::string("not a derived intensity view: ") + name);
    const size_t n = derived_size(c, name);
    double *tmp = static_cast<double *>(c.dalloc(n * 8));
    const InflowSlices bs{c.BV, c.BH, c.refl, c.sg.TX, c.sg.TY, c.sg.NTX, c.sg.NTY, rec_lanes(c.ng * c.sg.K)};
    if (c.precision)
        hipLaunchKernelGGL(derived_kernel<float>, dim3(1024), dim3(256), 0, c.stream,
                           static_cast<const float *>(c.Icrn[c.icur]), c.bcin, bs, tmp, kind, c.nx, c.ny, c.nm, c.ng);
    else
        hipLaunchKernelGGL(derived_kernel<double>, dim3(1024), dim3(256), 0, c.stream,
                           static_cast<const double *>(c.Icrn[c.icur]), c.bcin, bs, tmp, kind, c.nx, c.ny, c.nm, c.ng);
    TRT_CHECK(hipMemcpyAsync(host, tmp, n * 8, hipMemcpyDeviceToHost, c.stream));
    TRT_CHECK(hipStreamSynchronize(c.stream));
    c.dfree_last();
}

// (KapE + 1/(c dt), RT_Src) per (cell, group) for the sweep, from the named arrays
__global__ void sc_kernel(const double *KE, const double *RS, double2 *sc, float2 *scf, size_t nc, int ng,
                          double inv_cdt)
{
    const size_t n = nc * ng;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x) {
        const size_t cc = idx % nc;
        const int g = (int)(idx / nc);
        sc[cc * ng + g] = make_double2(KE[idx] + inv_cdt, RS[idx]);
        if (scf) scf[cc * ng + g] = make_float2((float)(KE[idx] + inv_cdt), (float)(RS[idx] * kIScale32));
    }
}

void rebuild_sc(Ctx &c)
{
    hipLaunchKernelGGL(sc_kernel, dim3(1024), dim3(256), 0, c.stream, c.arr("KapE"), c.arr("RT_Src"),
                       reinterpret_cast<double2 *>(c.sc), reinterpret_cast<float2 *>(c.scf), (size_t)c.nx * c.ny, c.ng,
                       1.0 / (c.p.c * c.p.dt));
    TRT_CHECK(hipGetLastError());
}

int sweep_launches(const Ctx &c)
{
    return c.sweep_mode == 2 || c.sweep_mode == 4 ? 1 : c.sg.NTX + c.sg.NTY - 1;
}

// plain device copy of n doubles, the calibration kernel for HBM byte counters
__global__ void calib_copy_kernel(const double2 *__restrict__ src, double2 *__restrict__ dst, size_t n2)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n2; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

void calib_copy(size_t nbytes, int reps, double *ms)
{
    const size_t n2 = nbytes / 16;
    void *a = nullptr, *b = nullptr;
    TRT_CHECK(hipMalloc(&a, n2 * 16));
    TRT_CHECK(hipMalloc(&b, n2 * 16));
    TRT_CHECK(hipMemset(a, 0, n2 * 16));
    hipEvent_t e0, e1;
    TRT_CHECK(hipEventCreate(&e0));
    TRT_CHECK(hipEventCreate(&e1));
    hipLaunchKernelGGL(calib_copy_kernel, dim3(4096), dim3(256), 0, nullptr, (const double2 *)a, (double2 *)b, n2);
    TRT_CHECK(hipEventRecord(e0, nullptr));
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(calib_copy_kernel, dim3(4096), dim3(256), 0, nullptr, (const double2 *)a, (double2 *)b, n2);
    TRT_CHECK(hipEventRecord(e1, nullptr));
    TRT_CHECK(hipEventSynchronize(e1));
    float t = 0.f;
    TRT_CHECK(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
}

}  // namespace trt
