// Low-order part of the MLQD iteration on the GPU.
//
// OLD_MGQD_COEFS  MainSrc/MLOQD_SOLVES.f08:12-73
// MLOQD_FV        MainSrc/MLOQD_SOLVES.f08:78-362 (coefficients :148-221, flux recovery :242-261)
// QD_FV           MainSrc/QD_SOLVERS.f08:40-538 (reduced E-only finite-volume system)
// OLD_GREY_COEFS / GREY_COEFS / Cbar_Calc / EGP_FV_NEWT  MainSrc/GLOQD_SOLVES.f08
// TRT_MLQD_ALGORITHM step  MainSrc/ALGORITHMS.f08:389-897
//
// The reduced QD system is kept matrix-free on a structured layout: per system
// the unknown vector is [E_avg (N_x,N_y) | E_edgV (N_x+1,N_y) | E_edgH (N_x,N_y+1)]
// and each row is the equation owned by that unknown, exactly the rows the
// reference assembles in CSR (QD_SOLVERS.f08:91-400): energy balance at a cell
// centre, x-momentum continuity (or the left/right boundary condition) at a
// vertical face, y-momentum continuity (or bottom/top condition) at a horizontal
// face.  All N_g group systems of MLOQD_FV are solved as one batch by a
// Jacobi-preconditioned BiCGSTAB; the grey Newton system is a batch of one.
#include "qd_view.h"
#include "physics_device.h"

#include <chrono>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>

namespace trt {

static int nb(size_t n) { size_t b = (n + 255) / 256; return (int)std::min<size_t>(std::max<size_t>(b, 1), 65536); }

// Row `row` of system `sys` applied to the scaled unknowns y (x = S y, S = diag(cs)):
// ax = (A S y)_row, diag = A_row,row * cs_row, rhs = b_row.
__device__ __forceinline__ void qd_row(const QDView &V, int sys, size_t row, const double *y, double &ax,
                                       double &diag, double &rhs)
{
    const double *s = V.cs + (size_t)sys * V.nrow;
    double acc = 0.0;
    qd_row_visit(V, sys, row, [&](int kind, int i, int j, double a) {
        const size_t k = qd_index(V, kind, i, j);
        acc += a * (y[k] * s[k]);
    }, diag, rhs);
    ax = acc;
    diag *= s[row];
}

// ---------------------------------------------------------------------------
// batched BiCGSTAB on the two-sided scaled system D^-1 A S y = D^-1 b, x = S y,
// S = diag(|x_ref| + floor), D = diag(A S).  The unknowns of one system span
// many orders of magnitude (E in a cold region is ~1e-12 of E near a hot
// boundary) and so do the rows, so the solve is posed in relative terms: the
// stopping test then bounds every unknown's error relative to its own size,
// which a norm of the unscaled residual cannot do.
// ---------------------------------------------------------------------------
// per-system scalars
enum { S_RHO, S_ALPHA, S_OMEGA, S_BETA, S_BNORM, S_RNORM, S_DONE, S_ITS, S_XMAX, S_ACC, S_RNORM0, S_NSCAL };

constexpr int RT = 256;   // threads of the vector kernels

__device__ __forceinline__ double block_sum(double v, double *sh)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) sh[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += sh[k];
    return t;   // valid on thread 0
}

struct KrylovArgs {
    QDView V;
    int nsys, RB;
    const double *b;
    double *x, *r, *rh, *p, *v, *s, *t, *dg;
    double *red;      // [sys][4][RB]
    double *scal;     // [sys][S_NSCAL]
    double tol;
};

#define SYS_LOOP(a)                                                                             \
    const int sys = blockIdx.y;                                                                 \
    if (a.scal[sys * S_NSCAL + S_DONE] != 0.0) return;                                          \
    const size_t off = (size_t)sys * a.V.nrow;                                                  \
    for (size_t row = (size_t)blockIdx.x * RT + threadIdx.x; row < a.V.nrow; row += (size_t)a.RB * RT)

// r = b - A x, rh = r, p = v = 0, diag; partial (b,b) and (r,r)
__global__ __launch_bounds__(RT) void k_init(KrylovArgs a)
{
    __shared__ double shm[8];
    double bb = 0.0, rr = 0.0;
    const int sys = blockIdx.y;
    const size_t off = (size_t)sys * a.V.nrow;
    for (size_t row = (size_t)blockIdx.x * RT + threadIdx.x; row < a.V.nrow; row += (size_t)a.RB * RT) {
        double ax, dg, rhs;
        qd_row(a.V, sys, row, a.x + off, ax, dg, rhs);
        const double r = (rhs - ax) / dg, bs = rhs / dg;
        a.r[off + row] = r; a.rh[off + row] = r; a.p[off + row] = 0.0; a.v[off + row] = 0.0;
        a.dg[off + row] = dg;
        bb += bs * bs; rr += r * r;
    }
    const double s1 = block_sum(bb, shm);
    const double s2 = block_sum(rr, shm);
    if (threadIdx.x == 0) {
        a.red[((size_t)sys * 4 + 0) * a.RB + blockIdx.x] = s1;
        a.red[((size_t)sys * 4 + 1) * a.RB + blockIdx.x] = s2;
    }
}

__global__ void k_init_fin(KrylovArgs a)
{
    const int sys = blockIdx.x * blockDim.x + threadIdx.x;
    if (sys >= a.nsys) return;
    double bb = 0.0, rr = 0.0;
    for (int k = 0; k < a.RB; ++k) { bb += a.red[((size_t)sys * 4 + 0) * a.RB + k]; rr += a.red[((size_t)sys * 4 + 1) * a.RB + k]; }
    double *S = a.scal + sys * S_NSCAL;
    S[S_BNORM] = sqrt(bb);
    S[S_RNORM] = sqrt(rr);
    S[S_RNORM0] = sqrt(rr);
    if (S[S_ACC] != 0.0) {   // accepted in an earlier stage: leave it alone
        S[S_DONE] = 1.0;
        S[S_ITS] = 0.0;
        return;
    }
    S[S_RHO] = rr;           // (rh, r) with rh = r
    S[S_ALPHA] = 1.0; S[S_OMEGA] = 1.0; S[S_BETA] = 0.0;
    S[S_ITS] = 0.0;
    S[S_DONE] = (sqrt(rr) <= a.tol * sqrt(bb) || rr == 0.0) ? 1.0 : 0.0;
}

// p = r + beta (p - omega v)
__global__ __launch_bounds__(RT) void k_p(KrylovArgs a)
{
    const double *S = a.scal + blockIdx.y * S_NSCAL;
    const double beta = S[S_BETA], om = S[S_OMEGA];
    SYS_LOOP(a) {
        const size_t k = off + row;
        const double p = a.r[k] + beta * (a.p[k] - om * a.v[k]);
        a.p[k] = p;
    }
}

// v = D^-1 A p; partial (rh, v)
__global__ __launch_bounds__(RT) void k_av(KrylovArgs a)
{
    __shared__ double shm[8];
    const int sys = blockIdx.y;
    if (a.scal[sys * S_NSCAL + S_DONE] != 0.0) return;
    const size_t off = (size_t)sys * a.V.nrow;
    double acc = 0.0;
    for (size_t row = (size_t)blockIdx.x * RT + threadIdx.x; row < a.V.nrow; row += (size_t)a.RB * RT) {
        double ax, dg, rhs;
        qd_row(a.V, sys, row, a.p + off, ax, dg, rhs);
        const double v = ax / dg;
        a.v[off + row] = v;
        acc += a.rh[off + row] * v;
    }
    const double s = block_sum(acc, shm);
    if (threadIdx.x == 0) a.red[((size_t)sys * 4 + 0) * a.RB + blockIdx.x] = s;
}

__global__ void k_alpha(KrylovArgs a)
{
    const int sys = blockIdx.x * blockDim.x + threadIdx.x;
    if (sys >= a.nsys) return;
    double *S = a.scal + sys * S_NSCAL;
    if (S[S_DONE] != 0.0) return;
    double d = 0.0;
    for (int k = 0; k < a.RB; ++k) d += a.red[((size_t)sys * 4 + 0) * a.RB + k];
    if (d == 0.0 || !isfinite(d)) { S[S_DONE] = 2.0; return; }   // breakdown
    S[S_ALPHA] = S[S_RHO] / d;
}

// s = r - alpha v
__global__ __launch_bounds__(RT) void k_s(KrylovArgs a)
{
    const double al = a.scal[blockIdx.y * S_NSCAL + S_ALPHA];
    SYS_LOOP(a) {
        const size_t k = off + row;
        const double s = a.r[k] - al * a.v[k];
        a.s[k] = s;
    }
}

// t = A sh; partials (t,s), (t,t)
__global__ __launch_bounds__(RT) void k_at(KrylovArgs a)
{
    __shared__ double shm[8];
    const int sys = blockIdx.y;
    if (a.scal[sys * S_NSCAL + S_DONE] != 0.0) return;
    const size_t off = (size_t)sys * a.V.nrow;
    double ts = 0.0, tt = 0.0;
    for (size_t row = (size_t)blockIdx.x * RT + threadIdx.x; row < a.V.nrow; row += (size_t)a.RB * RT) {
        double ax, dg, rhs;
        qd_row(a.V, sys, row, a.s + off, ax, dg, rhs);
        const double t = ax / dg;
        a.t[off + row] = t;
        ts += t * a.s[off + row];
        tt += t * t;
    }
    const double s1 = block_sum(ts, shm);
    const double s2 = block_sum(tt, shm);
    if (threadIdx.x == 0) {
        a.red[((size_t)sys * 4 + 0) * a.RB + blockIdx.x] = s1;
        a.red[((size_t)sys * 4 + 1) * a.RB + blockIdx.x] = s2;
    }
}

__global__ void k_omega(KrylovArgs a)
{
    const int sys = blockIdx.x * blockDim.x + threadIdx.x;
    if (sys >= a.nsys) return;
    double *S = a.scal + sys * S_NSCAL;
    if (S[S_DONE] != 0.0) return;
    double ts = 0.0, tt = 0.0;
    for (int k = 0; k < a.RB; ++k) { ts += a.red[((size_t)sys * 4 + 0) * a.RB + k]; tt += a.red[((size_t)sys * 4 + 1) * a.RB + k]; }
    S[S_OMEGA] = tt > 0.0 ? ts / tt : 0.0;
}

// x += alpha p + omega s; r = s - omega t; partials (r,r), (rh,r)
__global__ __launch_bounds__(RT) void k_xr(KrylovArgs a)
{
    __shared__ double shm[8];
    const int sys = blockIdx.y;
    const double *S = a.scal + sys * S_NSCAL;
    if (S[S_DONE] != 0.0) return;
    const double al = S[S_ALPHA], om = S[S_OMEGA];
    const size_t off = (size_t)sys * a.V.nrow;
    double rr = 0.0, rhr = 0.0;
    for (size_t row = (size_t)blockIdx.x * RT + threadIdx.x; row < a.V.nrow; row += (size_t)a.RB * RT) {
        const size_t k = off + row;
        a.x[k] += al * a.p[k] + om * a.s[k];
        const double r = a.s[k] - om * a.t[k];
        a.r[k] = r;
        rr += r * r;
        rhr += a.rh[k] * r;
    }
    const double s1 = block_sum(rr, shm);
    const double s2 = block_sum(rhr, shm);
    if (threadIdx.x == 0) {
        a.red[((size_t)sys * 4 + 0) * a.RB + blockIdx.x] = s1;
        a.red[((size_t)sys * 4 + 1) * a.RB + blockIdx.x] = s2;
    }
}

__global__ void k_rho(KrylovArgs a)
{
    const int sys = blockIdx.x * blockDim.x + threadIdx.x;
    if (sys >= a.nsys) return;
    double *S = a.scal + sys * S_NSCAL;
    if (S[S_DONE] != 0.0) return;
    double rr = 0.0, rhr = 0.0;
    for (int k = 0; k < a.RB; ++k) { rr += a.red[((size_t)sys * 4 + 0) * a.RB + k]; rhr += a.red[((size_t)sys * 4 + 1) * a.RB + k]; }
    S[S_ITS] += 1.0;
    S[S_RNORM] = sqrt(rr);
    if (sqrt(rr) <= a.tol * S[S_BNORM]) { S[S_DONE] = 1.0; return; }
    // diverging: the recurrence residual grew far beyond the initial one
    if (sqrt(rr) > 1e4 * fmax(S[S_RNORM0], a.tol * S[S_BNORM])) { S[S_DONE] = 2.0; return; }
    if (S[S_OMEGA] == 0.0 || S[S_RHO] == 0.0 || !isfinite(rr)) { S[S_DONE] = 2.0; return; }
    S[S_BETA] = (rhr / S[S_RHO]) * (S[S_ALPHA] / S[S_OMEGA]);
    S[S_RHO] = rhr;
}

// true residual norms, for the final acceptance test: |r| and the normwise
// backward-error scale | |A||x| + |b| |, both in the diagonally scaled rows
__global__ __launch_bounds__(RT) void k_true_res(KrylovArgs a)
{
    __shared__ double shm[8];
    const int sys = blockIdx.y;
    const size_t off = (size_t)sys * a.V.nrow;
    const double *cs = a.V.cs + off;
    const double *y = a.x + off;
    double rr = 0.0, ss = 0.0;
    for (size_t row = (size_t)blockIdx.x * RT + threadIdx.x; row < a.V.nrow; row += (size_t)a.RB * RT) {
        double ax = 0.0, aax = 0.0, dg, rhs;
        qd_row_visit(a.V, sys, row, [&](int kind, int i, int j, double c) {
            const double t = c * (y[qd_index(a.V, kind, i, j)] * cs[qd_index(a.V, kind, i, j)]);
            ax += t;
            aax += fabs(t);
        }, dg, rhs);
        dg = fabs(dg * cs[row]);
        const double r = (rhs - ax) / dg, sc = (aax + fabs(rhs)) / dg;
        rr += r * r;
        ss += sc * sc;
    }
    const double s1 = block_sum(rr, shm);
    const double s2 = block_sum(ss, shm);
    if (threadIdx.x == 0) {
        a.red[((size_t)sys * 4 + 2) * a.RB + blockIdx.x] = s1;
        a.red[((size_t)sys * 4 + 3) * a.RB + blockIdx.x] = s2;
    }
}

// true residual norms; a system whose residual meets thr * |b| is accepted.
// out[4 sys ..] = |r|, |b|, accepted, | |A||x| + |b| |
__global__ void k_true_fin(KrylovArgs a, double *out, double thr)
{
    const int sys = blockIdx.x * blockDim.x + threadIdx.x;
    if (sys >= a.nsys) return;
    double rr = 0.0, ss = 0.0;
    for (int k = 0; k < a.RB; ++k) {
        rr += a.red[((size_t)sys * 4 + 2) * a.RB + k];
        ss += a.red[((size_t)sys * 4 + 3) * a.RB + k];
    }
    double *S = a.scal + sys * S_NSCAL;
    const double r = sqrt(rr), b = S[S_BNORM];
    if (r <= thr * b || rr == 0.0) S[S_ACC] = 1.0;
    out[4 * sys] = r;
    out[4 * sys + 1] = b;
    out[4 * sys + 2] = S[S_ACC];
    out[4 * sys + 3] = sqrt(ss);
}

__global__ void k_clear_acc(KrylovArgs a)
{
    const int sys = blockIdx.x * blockDim.x + threadIdx.x;
    if (sys < a.nsys) a.scal[sys * S_NSCAL + S_ACC] = 0.0;
}

// systems not accepted by a stage start the next one from the original guess
__global__ __launch_bounds__(RT) void k_restore(KrylovArgs a, const double *x0)
{
    const int sys = blockIdx.y;
    if (a.scal[sys * S_NSCAL + S_ACC] != 0.0) return;
    const size_t off = (size_t)sys * a.V.nrow;
    for (size_t row = (size_t)blockIdx.x * RT + threadIdx.x; row < a.V.nrow; row += (size_t)a.RB * RT)
        a.x[off + row] = x0[off + row];
}

// column scales from a reference vector (the initial guess, or the state the
// unknowns correct): per-system max |ref| first, then cs = |ref| + 1e-20 max
__global__ __launch_bounds__(RT) void k_refmax(KrylovArgs a, const double *ref)
{
    __shared__ double shm[8];
    const int sys = blockIdx.y;
    const size_t off = (size_t)sys * a.V.nrow;
    double m = 0.0;
    for (size_t row = (size_t)blockIdx.x * RT + threadIdx.x; row < a.V.nrow; row += (size_t)a.RB * RT)
        m = fmax(m, fabs(ref[off + row]));
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) shm[w] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t = fmax(t, shm[k]);
        a.red[((size_t)sys * 4 + 3) * a.RB + blockIdx.x] = t;
    }
}

__global__ void k_refmax_fin(KrylovArgs a)
{
    const int sys = blockIdx.x * blockDim.x + threadIdx.x;
    if (sys >= a.nsys) return;
    double m = 0.0;
    for (int k = 0; k < a.RB; ++k) m = fmax(m, a.red[((size_t)sys * 4 + 3) * a.RB + k]);
    a.scal[sys * S_NSCAL + S_XMAX] = m;
}

// cs = |ref| + rel_floor * max|ref| (cs = 1 when rel_floor < 0); x (unscaled guess) -> y = x / cs
__global__ __launch_bounds__(RT) void k_colscale(KrylovArgs a, const double *ref, double *cs, double rel_floor)
{
    const int sys = blockIdx.y;
    const double mx = a.scal[sys * S_NSCAL + S_XMAX];
    const bool unit = rel_floor < 0.0 || !(mx > 0.0);
    const double floor = unit ? 1.0 : rel_floor * mx;
    const size_t off = (size_t)sys * a.V.nrow;
    for (size_t row = (size_t)blockIdx.x * RT + threadIdx.x; row < a.V.nrow; row += (size_t)a.RB * RT) {
        const size_t k = off + row;
        const double s = unit ? 1.0 : fabs(ref[k]) + floor;
        cs[k] = s;
        a.x[k] = a.x[k] / s;
    }
}

// x = S y
__global__ __launch_bounds__(RT) void k_unscale(KrylovArgs a)
{
    const int sys = blockIdx.y;
    const size_t off = (size_t)sys * a.V.nrow;
    for (size_t row = (size_t)blockIdx.x * RT + threadIdx.x; row < a.V.nrow; row += (size_t)a.RB * RT)
        a.x[off + row] *= a.V.cs[off + row];
}

void qd_alloc(Ctx &c, QDSystems &s, int nsys)
{
    s.nsys = nsys;
    s.nx = c.nx; s.ny = c.ny;
    const size_t nc = (size_t)c.nx * c.ny;
    s.nrow = 3 * nc + c.nx + c.ny;
    const size_t nbd = (size_t)std::max(c.nx, c.ny);
    s.coef = static_cast<double *>(c.dalloc((size_t)nsys * NCOEF * nc * 8));
    s.bnd = static_cast<double *>(c.dalloc((size_t)nsys * NBND * nbd * 8));
    s.b = nullptr;
    s.x = static_cast<double *>(c.dalloc((size_t)nsys * s.nrow * 8));
    s.diag = static_cast<double *>(c.dalloc((size_t)nsys * s.nrow * 8));
    s.cs = static_cast<double *>(c.dalloc((size_t)nsys * s.nrow * 8));
    s.ref = static_cast<double *>(c.dalloc((size_t)nsys * s.nrow * 8));
    s.x0 = static_cast<double *>(c.dalloc((size_t)nsys * s.nrow * 8));
    s.work = static_cast<double *>(c.dalloc((size_t)6 * nsys * s.nrow * 8));
    s.red_blocks = (int)std::min<size_t>(256, std::max<size_t>(1, (s.nrow + RT * 4 - 1) / (RT * 4)));
    s.red = static_cast<double *>(c.dalloc((size_t)nsys * 4 * s.red_blocks * 8));
    s.scal = static_cast<double *>(c.dalloc((size_t)nsys * S_NSCAL * 8));
    TRT_CHECK(hipMemsetAsync(s.coef, 0, (size_t)nsys * NCOEF * nc * 8, c.stream));
    TRT_CHECK(hipMemsetAsync(s.bnd, 0, (size_t)nsys * NBND * nbd * 8, c.stream));
    TRT_CHECK(hipMemsetAsync(s.x, 0, (size_t)nsys * s.nrow * 8, c.stream));
}

static KrylovArgs kargs(Ctx &c, QDSystems &s)
{
    KrylovArgs a{};
    a.V.nx = s.nx; a.V.ny = s.ny; a.V.nb = std::max(s.nx, s.ny);
    a.V.nc = (size_t)s.nx * s.ny; a.V.nv = (size_t)(s.nx + 1) * s.ny; a.V.nh = (size_t)s.nx * (s.ny + 1);
    a.V.nrow = s.nrow;
    a.V.coef = s.coef; a.V.bnd = s.bnd; a.V.cs = s.cs;
    a.nsys = s.nsys; a.RB = s.red_blocks;
    a.x = s.x;
    const size_t n = (size_t)s.nsys * s.nrow;
    double *w = s.work;
    a.r = w; a.rh = w + n; a.p = w + 2 * n; a.v = w + 3 * n; a.s = w + 4 * n; a.t = w + 5 * n;
    a.dg = s.diag;
    a.red = s.red; a.scal = s.scal;
    a.tol = c.lin_rtol;
    return a;
}

// Solves every system in up to three stages of decreasing column scaling:
// scales |ref| + 1e-20 max|ref| (relative accuracy of every unknown), then
// |ref| + 1e-6 max|ref|, then none.  Strong scaling can leave the scaled matrix
// too far from diagonal dominance for Jacobi-BiCGSTAB (a hot cell next to a
// cold one), so a system that a stage does not bring to 4 rtol on its true
// residual restarts the next stage from its original guess.
// True residual of the listed systems in diagonally scaled rows, |r| / |b|.
// Residual measures of the listed systems after a direct solve: |r|/|b| and
// the normwise backward error |r| / | |A||x| + |b| | (rows diagonally scaled).
struct DirectRes {
    std::vector<double> rel, bkwd;
};

static DirectRes direct_residuals(Ctx &c, QDSystems &s, KrylovArgs &a, const std::vector<int> &which)
{
    const dim3 grid(a.RB, s.nsys), blk(RT);
    const int fin_blocks = (s.nsys + 63) / 64;
    hipLaunchKernelGGL(k_colscale, grid, blk, 0, c.stream, a, s.ref, s.cs, -1.0);   // unit scales
    hipLaunchKernelGGL(k_init, grid, blk, 0, c.stream, a);                          // |b| in the unscaled rows
    hipLaunchKernelGGL(k_init_fin, dim3(fin_blocks), dim3(64), 0, c.stream, a);
    hipLaunchKernelGGL(k_true_res, grid, blk, 0, c.stream, a);
    hipLaunchKernelGGL(k_true_fin, dim3(fin_blocks), dim3(64), 0, c.stream, a, c.scratch, 1e4 * c.lin_rtol);
    std::vector<double> h(4 * (size_t)s.nsys);
    TRT_CHECK(hipMemcpyAsync(h.data(), c.scratch, h.size() * 8, hipMemcpyDeviceToHost, c.stream));
    TRT_CHECK(hipStreamSynchronize(c.stream));
    DirectRes d;
    for (int k : which) {
        // a zero residual is exact, also for b = 0 (x = 0)
        d.rel.push_back(h[4 * k] == 0.0 ? 0.0 : h[4 * k] / h[4 * k + 1]);
        d.bkwd.push_back(h[4 * k] == 0.0 ? 0.0 : h[4 * k] / h[4 * k + 3]);
    }
    return d;
}

// Direct solve of the listed systems by nested dissection.  Acceptance is on
// the normwise backward error |r| / | |A||x| + |b| | (rows diagonally scaled),
// which a stable elimination keeps near the rounding unit whatever the
// conditioning: |r|/|b| is no measure here, since the grey Newton systems of
// EGP_FV_NEWT have right-hand sides (the nonlinear residual) far smaller than
// |A||x| near convergence, and b = 0 exactly in groups whose Planck function
// underflows.  A system whose backward error is not finite or above 1e-11 -- a
// front whose restricted pivoting failed -- is solved again by block LU over
// grid rows, whose dense row blocks pivot over whole grid rows; one still above
// 1e-9 throws.  qd_solver 2 uses block LU for every system, under the same check.
// (The reference's direct fallback, GAUSS_COMPLETE, checks nothing.)
static void direct_and_check(Ctx &c, QDSystems &s, KrylovArgs &a, const std::vector<int> &which)
{
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<int> bad;
    int steps = -1;
    if (c.qd_solver == 2) {
        bad = which;       // block LU chosen for every system
    } else {
        try {
            // qd_solver 0: first the factors of these systems' last factorization in this
            // time step, by iterative refinement to the same backward error; a fresh
            // factorization when they no longer contract
            steps = c.qd_solver == 0 ? qd_nd_reuse(c, s, which, 10) : -1;
            if (steps >= 0) {
                c.nd_reused += 1;
                c.nd_reuse_steps += steps;
            } else {
                qd_nd(c, s, which);
                c.nd_factor += 1;
            }
            const DirectRes d = direct_residuals(c, s, a, which);
            for (size_t q = 0; q < which.size(); ++q)
                if (!(d.bkwd[q] <= 1e-11)) bad.push_back(which[q]);
        } catch (const Error &e) {
            if (e.msg.find("singular") == std::string::npos) throw;
            bad = which;
        }
        if (!bad.empty() && c.verbose) fprintf(stderr, "[qd_solve] %zu systems again by block LU\n", bad.size());
    }
    if (!bad.empty()) qd_direct(c, s, bad);
    const DirectRes d = direct_residuals(c, s, a, which);
    if (c.verbose > 1) {
        size_t w = 0;
        for (size_t q = 1; q < which.size(); ++q)
            if (!(d.bkwd[q] <= d.bkwd[w])) w = q;
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        fprintf(stderr, "[qd_solve] direct: %zu of %d systems of %zu rows in %.2f ms (%s %d), worst backward error "
                "%.3e (system %d, |r|/|b| = %.3e)\n", which.size(), s.nsys, s.nrow, ms,
                steps >= 0 ? "reused factors, refinement steps" : "factorized, qd_solver", steps >= 0 ? steps : c.qd_solver,
                d.bkwd[w], which[w], d.rel[w]);
    }
    for (size_t q = 0; q < which.size(); ++q)
        if (!(d.bkwd[q] <= 1e-9)) {
            // name non-finite coefficients: those come from the state, not the solve
            const int k = which[q];
            const size_t nc = (size_t)s.nx * s.ny, nb = (size_t)std::max(s.nx, s.ny);
            std::vector<double> cf(NCOEF * nc + NBND * nb);
            TRT_CHECK(hipMemcpyAsync(cf.data(), s.coef + (size_t)k * NCOEF * nc, NCOEF * nc * 8, hipMemcpyDeviceToHost,
                                     c.stream));
            TRT_CHECK(hipMemcpyAsync(cf.data() + NCOEF * nc, s.bnd + (size_t)k * NBND * nb, NBND * nb * 8,
                                     hipMemcpyDeviceToHost, c.stream));
            TRT_CHECK(hipStreamSynchronize(c.stream));
            size_t bad_cf = 0, bad_bd = 0;
            for (size_t t = 0; t < cf.size(); ++t)
                if (!std::isfinite(cf[t])) ++(t < NCOEF * nc ? bad_cf : bad_bd);
            char msg[260];
            snprintf(msg, sizeof msg,
                     "QD direct solve inaccurate (system %d of %d, %zu rows, backward error %.3e, |r|/|b| = %.3e; "
                     "%zu non-finite coefficients, %zu non-finite boundary terms)",
                     k, s.nsys, s.nrow, d.bkwd[q], d.rel[q], bad_cf, bad_bd);
            throw Error(msg);
        }
}

void qd_solve(Ctx &c, QDSystems &s, int *its_out, bool ref_given)
{
    KrylovArgs a = kargs(c, s);
    const dim3 grid(a.RB, s.nsys), blk(RT);
    const int fin_blocks = (s.nsys + 63) / 64;
    const size_t n = (size_t)s.nsys * s.nrow;
    if (c.qd_solver != 1) {
        // every system by the direct solver (reported as -1 iterations, the reference's
        // flag for a direct solve, QS:499)
        std::vector<int> all(s.nsys);
        for (int k = 0; k < s.nsys; ++k) all[k] = k;
        hipLaunchKernelGGL(k_clear_acc, dim3(fin_blocks), dim3(64), 0, c.stream, a);
        direct_and_check(c, s, a, all);
        if (its_out) for (int k = 0; k < s.nsys; ++k) its_out[k] = -1;
        return;
    }
    if (!ref_given) dcopy(c, s.ref, s.x, n);
    dcopy(c, s.x0, s.x, n);
    hipLaunchKernelGGL(k_refmax, grid, blk, 0, c.stream, a, s.ref);
    hipLaunchKernelGGL(k_refmax_fin, dim3(fin_blocks), dim3(64), 0, c.stream, a);
    hipLaunchKernelGGL(k_clear_acc, dim3(fin_blocks), dim3(64), 0, c.stream, a);
    // systems on which BiCGSTAB failed recently skip it: marked accepted, solved directly below
    if (s.direct_next.size() != (size_t)s.nsys) s.direct_next.assign(s.nsys, 0);
    std::vector<int> skipped;
    static const double one = 1.0;
    for (int k = 0; k < s.nsys; ++k)
        if (s.direct_next[k] > 0) {
            skipped.push_back(k);
            --s.direct_next[k];
            TRT_CHECK(hipMemcpyAsync(a.scal + (size_t)k * S_NSCAL + S_ACC, &one, 8, hipMemcpyHostToDevice, c.stream));
        }
    std::vector<double> hs((size_t)s.nsys * S_NSCAL);
    std::vector<int> total_its(s.nsys, 0);
    const double floors[3] = {1e-20, 1e-6, -1.0};
    bool all_ok = false;
    std::vector<double> h(4 * (size_t)s.nsys);
    for (int stage = 0; stage < 3 && !all_ok; ++stage) {
        if (skipped.size() == (size_t)s.nsys) break;
        if (stage > 0) hipLaunchKernelGGL(k_restore, grid, blk, 0, c.stream, a, s.x0);
        hipLaunchKernelGGL(k_colscale, grid, blk, 0, c.stream, a, s.ref, s.cs, floors[stage]);
        const double thr = (stage == 2 ? 100.0 : 4.0) * c.lin_rtol;
        for (int restart = 0; restart < 4; ++restart) {
            hipLaunchKernelGGL(k_init, grid, blk, 0, c.stream, a);
            hipLaunchKernelGGL(k_init_fin, dim3(fin_blocks), dim3(64), 0, c.stream, a);
            int it = 0;
            int check = 4;
            while (it < c.lin_maxit) {
                for (int k = 0; k < check && it < c.lin_maxit; ++k, ++it) {
                    hipLaunchKernelGGL(k_p, grid, blk, 0, c.stream, a);
                    hipLaunchKernelGGL(k_av, grid, blk, 0, c.stream, a);
                    hipLaunchKernelGGL(k_alpha, dim3(fin_blocks), dim3(64), 0, c.stream, a);
                    hipLaunchKernelGGL(k_s, grid, blk, 0, c.stream, a);
                    hipLaunchKernelGGL(k_at, grid, blk, 0, c.stream, a);
                    hipLaunchKernelGGL(k_omega, dim3(fin_blocks), dim3(64), 0, c.stream, a);
                    hipLaunchKernelGGL(k_xr, grid, blk, 0, c.stream, a);
                    hipLaunchKernelGGL(k_rho, dim3(fin_blocks), dim3(64), 0, c.stream, a);
                }
                TRT_CHECK(hipGetLastError());
                TRT_CHECK(hipMemcpyAsync(hs.data(), a.scal, hs.size() * 8, hipMemcpyDeviceToHost, c.stream));
                TRT_CHECK(hipStreamSynchronize(c.stream));
                bool all = true;
                for (int k = 0; k < s.nsys; ++k) all = all && hs[(size_t)k * S_NSCAL + S_DONE] != 0.0;
                if (all) break;
                check = std::min(check * 2, 64);
            }
            for (int k = 0; k < s.nsys; ++k) total_its[k] += (int)hs[(size_t)k * S_NSCAL + S_ITS];
            // acceptance on the true residual
            double *tr = c.scratch;
            hipLaunchKernelGGL(k_true_res, grid, blk, 0, c.stream, a);
            hipLaunchKernelGGL(k_true_fin, dim3(fin_blocks), dim3(64), 0, c.stream, a, tr, thr);
            TRT_CHECK(hipMemcpyAsync(h.data(), tr, h.size() * 8, hipMemcpyDeviceToHost, c.stream));
            TRT_CHECK(hipStreamSynchronize(c.stream));
            all_ok = true;
            bool diverged = false;
            for (int k = 0; k < s.nsys; ++k) {
                if (h[4 * k + 2] == 0.0) all_ok = false;
                if (!(h[4 * k] <= h[4 * k + 1])) diverged = true;    // worse than x = 0, or not finite
            }
            if (c.verbose > 1 || (c.verbose && !all_ok)) {
                int nacc = 0, worst = 0;
                for (int k = 0; k < s.nsys; ++k) {
                    nacc += h[4 * k + 2] != 0.0;
                    if (h[4 * k] / (h[4 * k + 1] + 1e-300) > h[4 * worst] / (h[4 * worst + 1] + 1e-300)) worst = k;
                }
                fprintf(stderr, "[qd_solve nsys=%d] stage %d restart %d its %d accepted %d/%d worst sys %d |r|/|b| %.3e\n",
                        s.nsys, stage, restart, it, nacc, s.nsys, worst, h[4 * worst] / (h[4 * worst + 1] + 1e-300));
            }
            if (all_ok || (diverged && stage < 2)) break;
        }
        hipLaunchKernelGGL(k_unscale, grid, blk, 0, c.stream, a);
    }
    TRT_CHECK(hipGetLastError());
    std::vector<int> which = skipped;
    for (int k = 0; k < s.nsys; ++k)
        if (h[4 * k + 2] == 0.0 && !std::binary_search(skipped.begin(), skipped.end(), k)) {
            which.push_back(k);
            s.direct_next[k] = 3;
        }
    if (!which.empty()) {
        // the reference falls back to a direct solve when BiCGSTAB fails (QS:492-500)
        std::sort(which.begin(), which.end());
        hipLaunchKernelGGL(k_restore, grid, blk, 0, c.stream, a, s.x0);
        hipLaunchKernelGGL(k_colscale, grid, blk, 0, c.stream, a, s.ref, s.cs, -1.0);   // unit scales
        hipLaunchKernelGGL(k_unscale, grid, blk, 0, c.stream, a);
        const auto t0 = std::chrono::steady_clock::now();
        direct_and_check(c, s, a, which);
        if (c.verbose) {
            const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            fprintf(stderr, "[qd_solve nsys=%d] direct solve of %zu systems (%zu skipped BiCGSTAB): %.1f ms\n", s.nsys,
                    which.size(), skipped.size(), ms);
        }
        for (int k : which) total_its[k] = -1;     // as the reference flags a direct solve (QS:499)
    }
    if (its_out) for (int k = 0; k < s.nsys; ++k) its_out[k] = total_its[k];
}

// gather E arrays into / scatter the solution out of the Krylov vector
__global__ void qd_gather(double *x, const double *Ea, const double *Ev, const double *Eh, size_t nc, size_t nv,
                          size_t nh, int nsys)
{
    const size_t nrow = nc + nv + nh, tot = nrow * nsys;
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < tot; k += (size_t)gridDim.x * blockDim.x) {
        const size_t s = k / nrow, r = k % nrow;
        x[k] = r < nc ? Ea[s * nc + r] : r < nc + nv ? Ev[s * nv + r - nc] : Eh[s * nh + r - nc - nv];
    }
}

__global__ void qd_scatter(const double *x, double *Ea, double *Ev, double *Eh, size_t nc, size_t nv, size_t nh,
                           int nsys)
{
    const size_t nrow = nc + nv + nh, tot = nrow * nsys;
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < tot; k += (size_t)gridDim.x * blockDim.x) {
        const size_t s = k / nrow, r = k % nrow;
        if (r < nc) Ea[s * nc + r] = x[k];
        else if (r < nc + nv) Ev[s * nv + r - nc] = x[k];
        else Eh[s * nh + r - nc - nv] = x[k];
    }
}

// ---------------------------------------------------------------------------
// OLD_MGQD_COEFS, MLOQD_SOLVES.f08:38-69
// ---------------------------------------------------------------------------
struct OldArgs {
    int nx, ny, ng;
    double c, f;
    const double *dx, *dy, *A;
    const double *Ea, *Ev, *Eh, *Fx, *Fy, *fxx, *fyy, *fvxx, *fvxy, *fhyy, *fhxy, *KE, *KR, *Src;
    double *G, *PL, *PB, *PR, *PT;
};

__global__ void old_mgqd_kernel(OldArgs a)
{
    const size_t nc = (size_t)a.nx * a.ny, tot = nc * a.ng;
    const size_t nv = (size_t)(a.nx + 1) * a.ny, nh = (size_t)a.nx * (a.ny + 1);
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < tot; k += (size_t)gridDim.x * blockDim.x) {
        const int g = (int)(k / nc);
        const size_t cc = k % nc;
        const int i = (int)(cc % a.nx), j = (int)(cc / a.nx);
        const size_t vL = g * nv + i + (size_t)(a.nx + 1) * j, vR = vL + 1;
        const size_t hB = g * nh + i + (size_t)a.nx * j, hT = hB + a.nx;
        const double dx = a.dx[i], dy = a.dy[j], A = a.A[cc], c = a.c, f = a.f;
        a.G[k] = f * (dy * (a.Fx[vR] - a.Fx[vL]) + dx * (a.Fy[hT] - a.Fy[hB]) + A * (c * a.KE[k] * a.Ea[k] - a.Src[k]));
        a.PL[k] = f * (c * dy * (a.fxx[k] * a.Ea[k] - a.fvxx[vL] * a.Ev[vL])
                       + c * dx * (a.fhxy[hT] * a.Eh[hT] - a.fhxy[hB] * a.Eh[hB]) / 2.0 + A * (a.KR[k] * a.Fx[vL]) / 2.0);
        a.PB[k] = f * (c * dx * (a.fyy[k] * a.Ea[k] - a.fhyy[hB] * a.Eh[hB])
                       + c * dy * (a.fvxy[vR] * a.Ev[vR] - a.fvxy[vL] * a.Ev[vL]) / 2.0 + A * (a.KR[k] * a.Fy[hB]) / 2.0);
        a.PR[k] = f * (c * dy * (a.fvxx[vR] * a.Ev[vR] - a.fxx[k] * a.Ea[k])
                       + c * dx * (a.fhxy[hT] * a.Eh[hT] - a.fhxy[hB] * a.Eh[hB]) / 2.0 + A * (a.KR[k] * a.Fx[vR]) / 2.0);
        a.PT[k] = f * (c * dx * (a.fhyy[hT] * a.Eh[hT] - a.fyy[k] * a.Ea[k])
                       + c * dy * (a.fvxy[vR] * a.Ev[vR] - a.fvxy[vL] * a.Ev[vL]) / 2.0 + A * (a.KR[k] * a.Fy[hT]) / 2.0);
    }
}

void old_mgqd_coefs(Ctx &c)
{
    OldArgs a{};
    a.nx = c.nx; a.ny = c.ny; a.ng = c.ng;
    a.c = c.p.c; a.f = (c.p.theta - 1.0) / c.p.theta;
    a.dx = c.dx; a.dy = c.dy; a.A = c.A;
    a.Ea = c.arr("Eg_avg_old"); a.Ev = c.arr("Eg_edgV_old"); a.Eh = c.arr("Eg_edgH_old");
    a.Fx = c.arr("Fxg_edgV_old"); a.Fy = c.arr("Fyg_edgH_old");
    a.fxx = c.arr("fg_avg_xx_old"); a.fyy = c.arr("fg_avg_yy_old");
    a.fvxx = c.arr("fg_edgV_xx_old"); a.fvxy = c.arr("fg_edgV_xy_old");
    a.fhyy = c.arr("fg_edgH_yy_old"); a.fhxy = c.arr("fg_edgH_xy_old");
    a.KE = c.arr("KapE_old"); a.KR = c.arr("KapR_old"); a.Src = c.arr("MGQD_Src_old");
    a.G = c.arr("G_old"); a.PL = c.arr("Pold_L"); a.PB = c.arr("Pold_B"); a.PR = c.arr("Pold_R"); a.PT = c.arr("Pold_T");
    hipLaunchKernelGGL(old_mgqd_kernel, dim3(nb((size_t)c.nx * c.ny * c.ng)), dim3(256), 0, c.stream, a);
    TRT_CHECK(hipGetLastError());
}

// OLD_GREY_COEFS, GLOQD_SOLVES.f08:24-39
__global__ void old_grey_kernel(int nx, int ny, int ng, double c, double th, double dt, double cv, const double *G,
                                const double *A, const double *Eo, const double *To, const double *KBo,
                                const double *Qo, double *Gh, double *Ro)
{
    const size_t nc = (size_t)nx * ny;
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < nc; k += (size_t)gridDim.x * blockDim.x) {
        double t = 0.0;
        for (int g = 0; g < ng; ++g) t += G[k + g * nc];
        Gh[k] = t + A[k] * Eo[k] / (th * dt);
        Ro[k] = cv / (th * dt) * To[k] + ((1.0 - th) / th) * (c * KBo[k] * Eo[k] - Qo[k]);
    }
}

void old_grey_coefs(Ctx &c)
{
    hipLaunchKernelGGL(old_grey_kernel, dim3(nb((size_t)c.nx * c.ny)), dim3(256), 0, c.stream, c.nx, c.ny, c.ng, c.p.c,
                       c.p.theta, c.p.dt, c.p.cv, c.arr("G_old"), c.A, c.arr("E_avg_old"), c.arr("Temp_Old"),
                       c.arr("KapE_Bar_old"), c.arr("GQD_Src_old"), c.arr("Gold_Hat"), c.arr("Rhat_old"));
    TRT_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// MLOQD_FV coefficients, MLOQD_SOLVES.f08:148-221
// ---------------------------------------------------------------------------
struct MLArgs {
    int nx, ny, ng, nbd;
    double c, dt, th, Fdt, Fin, Ein;
    const double *dx, *dy, *A;
    const double *fxx, *fyy, *fvxx, *fvxy, *fhyy, *fhxy, *KE, *KR, *Src, *Eao, *Fxo, *Fyo, *G, *PoL, *PoB, *PoR, *PoT;
    const double *CgL, *CgB, *CgR, *CgT, *EinL, *EinB, *EinR, *EinT, *FinL, *FinB, *FinR, *FinT;
    double *Xi, *PhL, *PhB, *PhR, *PhT;   // [g][nc]
    double *coef, *bnd;
};

// phase 1: Xi and the Phat terms (ML:148-165)
__global__ void ml_phase1(MLArgs a)
{
    const size_t nc = (size_t)a.nx * a.ny, tot = nc * a.ng;
    const size_t nv = (size_t)(a.nx + 1) * a.ny, nh = (size_t)a.nx * (a.ny + 1);
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < tot; k += (size_t)gridDim.x * blockDim.x) {
        const int g = (int)(k / nc);
        const size_t cc = k % nc;
        const int i = (int)(cc % a.nx), j = (int)(cc / a.nx);
        const double A = a.A[cc], c = a.c, th = a.th, dt = a.dt, Fdt = a.Fdt;
        const double Xi = A * (Fdt / (th * c * dt) + a.KR[k]) / 2.0;
        const size_t vL = g * nv + i + (size_t)(a.nx + 1) * j, hB = g * nh + i + (size_t)a.nx * j;
        const double PhL = (Fdt * A * a.Fxo[vL] / (2.0 * th * c * dt) + a.PoL[k]) / Xi;
        const double PhB = (Fdt * A * a.Fyo[hB] / (2.0 * th * c * dt) + a.PoB[k]) / Xi;
        const double PhR = (Fdt * A * a.Fxo[vL + 1] / (2.0 * th * c * dt) + a.PoR[k]) / Xi;
        const double PhT = (Fdt * A * a.Fyo[hB + a.nx] / (2.0 * th * c * dt) + a.PoT[k]) / Xi;
        a.Xi[k] = Xi; a.PhL[k] = PhL; a.PhB[k] = PhB; a.PhR[k] = PhR; a.PhT[k] = PhT;
        double *cf = a.coef + (size_t)g * NCOEF * nc;
        cf[ERHS * nc + cc] = A * a.Src[k] + A * a.Eao[k] / (th * dt) + a.G[k] + a.dy[j] * (PhL - PhR) + a.dx[i] * (PhB - PhT);
    }
}

// phase 2: EB, MBx, MBy coefficients and momentum right-hand sides (ML:167-196, 216-221)
__global__ void ml_phase2(MLArgs a)
{
    const size_t nc = (size_t)a.nx * a.ny, tot = nc * a.ng;
    const size_t nv = (size_t)(a.nx + 1) * a.ny, nh = (size_t)a.nx * (a.ny + 1);
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < tot; k += (size_t)gridDim.x * blockDim.x) {
        const int g = (int)(k / nc);
        const size_t cc = k % nc;
        const int i = (int)(cc % a.nx), j = (int)(cc / a.nx);
        const double A = a.A[cc], c = a.c, dx = a.dx[i], dy = a.dy[j];
        const double Xi = a.Xi[k];
        const size_t vL = g * nv + i + (size_t)(a.nx + 1) * j, vR = vL + 1;
        const size_t hB = g * nh + i + (size_t)a.nx * j, hT = hB + a.nx;
        double *cf = a.coef + (size_t)g * NCOEF * nc;
        cf[EBC * nc + cc] = (A * (1.0 / (a.th * a.dt) + c * a.KE[k]) + 2.0 * c * dy * dy * a.fxx[k] / Xi
                             + 2.0 * c * dx * dx * a.fyy[k] / Xi);
        cf[EBR * nc + cc] = -c * dy * dy * a.fvxx[vR] / Xi;
        cf[EBL * nc + cc] = -c * dy * dy * a.fvxx[vL] / Xi;
        cf[EBT * nc + cc] = -c * dx * dx * a.fhyy[hT] / Xi;
        cf[EBB * nc + cc] = -c * dx * dx * a.fhyy[hB] / Xi;
        cf[MXC * nc + cc] = c * dy * a.fxx[k] / Xi;
        cf[MXT * nc + cc] = -c * dx * a.fhxy[hT] / (2.0 * Xi);
        cf[MXB * nc + cc] = c * dx * a.fhxy[hB] / (2.0 * Xi);
        cf[MYC * nc + cc] = c * dx * a.fyy[k] / Xi;
        cf[MYR * nc + cc] = -c * dy * a.fvxy[vR] / (2.0 * Xi);
        cf[MYL * nc + cc] = c * dy * a.fvxy[vL] / (2.0 * Xi);
        if (i < a.nx - 1) {
            cf[MXR * nc + cc] = -(c * dy / Xi + c * dy / a.Xi[k + 1]) * a.fvxx[vR];
            cf[MXRHS * nc + cc] = a.PhL[k + 1] - a.PhR[k];
        }
        if (j < a.ny - 1) {
            cf[MYT * nc + cc] = -(c * dx / Xi + c * dx / a.Xi[k + a.nx]) * a.fhyy[hT];
            cf[MYRHS * nc + cc] = a.PhB[k + a.nx] - a.PhT[k];
        }
    }
}

// phase 3: boundary coefficients (ML:198-214)
__global__ void ml_phase3(MLArgs a)
{
    const size_t nc = (size_t)a.nx * a.ny;
    const size_t nv = (size_t)(a.nx + 1) * a.ny, nh = (size_t)a.nx * (a.ny + 1);
    const int per = 2 * a.ny + 2 * a.nx;
    const int tot = per * a.ng;
    const double c = a.c;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < tot; k += gridDim.x * blockDim.x) {
        const int g = k / per;
        int r = k % per;
        double *bd = a.bnd + (size_t)g * NBND * a.nbd;
        const double *Xi = a.Xi + g * nc;
        if (r < a.ny) {
            const int j = r;
            const size_t cg = j + (size_t)a.ny * g;
            bd[CPL * a.nbd + j] = c * a.CgL[cg] - c * a.dy[j] * a.fvxx[g * nv + (size_t)(a.nx + 1) * j] / Xi[(size_t)a.nx * j];
            bd[BCL * a.nbd + j] = a.Ein * c * a.CgL[cg] * a.EinL[cg] - a.Fin * a.FinL[cg] + a.PhL[g * nc + (size_t)a.nx * j];
        } else if (r < 2 * a.ny) {
            const int j = r - a.ny;
            const size_t cg = j + (size_t)a.ny * g;
            const size_t kc = (size_t)(a.nx - 1) + (size_t)a.nx * j;
            bd[CPR * a.nbd + j] = c * a.CgR[cg] + c * a.dy[j] * a.fvxx[g * nv + a.nx + (size_t)(a.nx + 1) * j] / Xi[kc];
            bd[BCR * a.nbd + j] = a.Ein * c * a.CgR[cg] * a.EinR[cg] - a.Fin * a.FinR[cg] + a.PhR[g * nc + kc];
        } else if (r < 2 * a.ny + a.nx) {
            const int i = r - 2 * a.ny;
            const size_t cg = i + (size_t)a.nx * g;
            bd[CPB * a.nbd + i] = c * a.CgB[cg] - c * a.dx[i] * a.fhyy[g * nh + i] / Xi[i];
            bd[BCB * a.nbd + i] = a.Ein * c * a.CgB[cg] * a.EinB[cg] - a.Fin * a.FinB[cg] + a.PhB[g * nc + i];
        } else {
            const int i = r - 2 * a.ny - a.nx;
            const size_t cg = i + (size_t)a.nx * g;
            const size_t kc = (size_t)i + (size_t)a.nx * (a.ny - 1);
            bd[CPT * a.nbd + i] = c * a.CgT[cg] + c * a.dx[i] * a.fhyy[g * nh + i + (size_t)a.nx * a.ny] / Xi[kc];
            bd[BCT * a.nbd + i] = a.Ein * c * a.CgT[cg] * a.EinT[cg] - a.Fin * a.FinT[cg] + a.PhT[g * nc + kc];
        }
    }
}

// flux recovery, ML:242-261
__global__ void ml_flux(MLArgs a, const double *Ea, const double *Ev, const double *Eh, double *Fx, double *Fy)
{
    const size_t nc = (size_t)a.nx * a.ny, tot = nc * a.ng;
    const size_t nv = (size_t)(a.nx + 1) * a.ny, nh = (size_t)a.nx * (a.ny + 1);
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < tot; k += (size_t)gridDim.x * blockDim.x) {
        const int g = (int)(k / nc);
        const size_t cc = k % nc;
        const int i = (int)(cc % a.nx), j = (int)(cc / a.nx);
        const double c = a.c, dx = a.dx[i], dy = a.dy[j], Xi = a.Xi[k];
        const size_t vL = g * nv + i + (size_t)(a.nx + 1) * j, vR = vL + 1;
        const size_t hB = g * nh + i + (size_t)a.nx * j, hT = hB + a.nx;
        Fx[vL] = -(c * dy / (Xi)) * (a.fxx[k] * Ea[k] - a.fvxx[vL] * Ev[vL])
                 - (c * dx / (2.0 * Xi)) * (a.fhxy[hT] * Eh[hT] - a.fhxy[hB] * Eh[hB]) + a.PhL[k];
        Fy[hB] = -(c * dx / (Xi)) * (a.fyy[k] * Ea[k] - a.fhyy[hB] * Eh[hB])
                 - (c * dy / (2.0 * Xi)) * (a.fvxy[vR] * Ev[vR] - a.fvxy[vL] * Ev[vL]) + a.PhB[k];
        if (i == a.nx - 1)
            Fx[vR] = -(c * dy / (Xi)) * (a.fvxx[vR] * Ev[vR] - a.fxx[k] * Ea[k])
                     - (c * dx / (2.0 * Xi)) * (a.fhxy[hT] * Eh[hT] - a.fhxy[hB] * Eh[hB]) + a.PhR[k];
        if (j == a.ny - 1)
            Fy[hT] = -(c * dx / (Xi)) * (a.fhyy[hT] * Eh[hT] - a.fyy[k] * Ea[k])
                     - (c * dy / (2.0 * Xi)) * (a.fvxy[vR] * Ev[vR] - a.fvxy[vL] * Ev[vL]) + a.PhT[k];
    }
}

static double *ml_scratch(Ctx &c)
{
    auto it = c.arrays.find("_ml_scratch");
    if (it != c.arrays.end()) return it->second.p;
    return c.reg("_ml_scratch", c.nx * c.ny, c.ng, 5).p;
}

void mloqd_fv(Ctx &c, int *kits)
{
    if (c.mg.nsys == 0) qd_alloc(c, c.mg, c.ng);
    const size_t nc = (size_t)c.nx * c.ny;
    double *scr = ml_scratch(c);
    MLArgs a{};
    a.nx = c.nx; a.ny = c.ny; a.ng = c.ng; a.nbd = std::max(c.nx, c.ny);
    a.c = c.p.c; a.dt = c.p.dt; a.th = c.p.theta; a.Fdt = 1.0; a.Fin = 1.0; a.Ein = 1.0;   // mlqd weights, AL:305-308
    a.dx = c.dx; a.dy = c.dy; a.A = c.A;
    a.fxx = c.arr("fg_avg_xx"); a.fyy = c.arr("fg_avg_yy"); a.fvxx = c.arr("fg_edgV_xx"); a.fvxy = c.arr("fg_edgV_xy");
    a.fhyy = c.arr("fg_edgH_yy"); a.fhxy = c.arr("fg_edgH_xy");
    a.KE = c.arr("KapE"); a.KR = c.arr("KapR"); a.Src = c.arr("MGQD_Src"); a.Eao = c.arr("Eg_avg_old");
    a.Fxo = c.arr("Fxg_edgV_old"); a.Fyo = c.arr("Fyg_edgH_old"); a.G = c.arr("G_old");
    a.PoL = c.arr("Pold_L"); a.PoB = c.arr("Pold_B"); a.PoR = c.arr("Pold_R"); a.PoT = c.arr("Pold_T");
    a.CgL = c.arr("Cg_L"); a.CgB = c.arr("Cg_B"); a.CgR = c.arr("Cg_R"); a.CgT = c.arr("Cg_T");
    a.EinL = c.arr("Eg_in_L"); a.EinB = c.arr("Eg_in_B"); a.EinR = c.arr("Eg_in_R"); a.EinT = c.arr("Eg_in_T");
    a.FinL = c.arr("Fg_in_L"); a.FinB = c.arr("Fg_in_B"); a.FinR = c.arr("Fg_in_R"); a.FinT = c.arr("Fg_in_T");
    a.Xi = scr; a.PhL = scr + nc * c.ng; a.PhB = scr + 2 * nc * c.ng; a.PhR = scr + 3 * nc * c.ng; a.PhT = scr + 4 * nc * c.ng;
    a.coef = c.mg.coef; a.bnd = c.mg.bnd;
    const size_t tot = nc * c.ng;
    hipLaunchKernelGGL(ml_phase1, dim3(nb(tot)), dim3(256), 0, c.stream, a);
    hipLaunchKernelGGL(ml_phase2, dim3(nb(tot)), dim3(256), 0, c.stream, a);
    hipLaunchKernelGGL(ml_phase3, dim3(nb((size_t)(2 * c.nx + 2 * c.ny) * c.ng)), dim3(256), 0, c.stream, a);
    double *Ea = c.arr("Eg_avg"), *Ev = c.arr("Eg_edgV"), *Eh = c.arr("Eg_edgH");
    const size_t nv = (size_t)(c.nx + 1) * c.ny, nh = (size_t)c.nx * (c.ny + 1);
    // current E_g are the initial guess (ML:228-230)
    hipLaunchKernelGGL(qd_gather, dim3(nb(c.mg.nrow * c.ng)), dim3(256), 0, c.stream, c.mg.x, Ea, Ev, Eh, nc, nv, nh, c.ng);
    TRT_CHECK(hipGetLastError());
    std::vector<int> its(c.ng);
    qd_solve(c, c.mg, its.data(), false);
    hipLaunchKernelGGL(qd_scatter, dim3(nb(c.mg.nrow * c.ng)), dim3(256), 0, c.stream, c.mg.x, Ea, Ev, Eh, nc, nv, nh, c.ng);
    hipLaunchKernelGGL(ml_flux, dim3(nb(tot)), dim3(256), 0, c.stream, a, Ea, Ev, Eh, c.arr("Fxg_edgV"), c.arr("Fyg_edgH"));
    TRT_CHECK(hipGetLastError());
    c.last_kits = its;
    if (kits) for (int g = 0; g < c.ng; ++g) kits[g] = its[g];
}

// ---------------------------------------------------------------------------
// Cbar_Calc, GLOQD_SOLVES.f08:191-229
// ---------------------------------------------------------------------------
__global__ void cbar_kernel(int nx, int ny, int ng, const double *CgL, const double *CgB, const double *CgR,
                            const double *CgT, const double *Ev, const double *Eh, const double *EinL,
                            const double *EinB, const double *EinR, const double *EinT, double *CbL, double *CbB,
                            double *CbR, double *CbT)
{
    const size_t nv = (size_t)(nx + 1) * ny, nh = (size_t)nx * (ny + 1);
    const int tot = ny + nx;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < tot; k += gridDim.x * blockDim.x) {
        if (k < ny) {
            const int j = k;
            double cl = 0, cr = 0, s1 = 0, s2 = 0;
            for (int g = 0; g < ng; ++g) {
                double q = Ev[g * nv + (size_t)(nx + 1) * j] - EinL[j + (size_t)ny * g];
                cl += CgL[j + (size_t)ny * g] * q; s1 += q;
                q = Ev[g * nv + nx + (size_t)(nx + 1) * j] - EinR[j + (size_t)ny * g];
                cr += CgR[j + (size_t)ny * g] * q; s2 += q;
            }
            CbL[j] = cl / s1; CbR[j] = cr / s2;
        } else {
            const int i = k - ny;
            double cb = 0, ct = 0, s1 = 0, s2 = 0;
            for (int g = 0; g < ng; ++g) {
                double q = Eh[g * nh + i] - EinB[i + (size_t)nx * g];
                cb += CgB[i + (size_t)nx * g] * q; s1 += q;
                q = Eh[g * nh + i + (size_t)nx * ny] - EinT[i + (size_t)nx * g];
                ct += CgT[i + (size_t)nx * g] * q; s2 += q;
            }
            CbB[i] = cb / s1; CbT[i] = ct / s2;
        }
    }
}

void cbar_calc(Ctx &c)
{
    hipLaunchKernelGGL(cbar_kernel, dim3(nb((size_t)c.nx + c.ny)), dim3(256), 0, c.stream, c.nx, c.ny, c.ng,
                       c.arr("Cg_L"), c.arr("Cg_B"), c.arr("Cg_R"), c.arr("Cg_T"), c.arr("Eg_edgV"), c.arr("Eg_edgH"),
                       c.arr("Eg_in_L"), c.arr("Eg_in_B"), c.arr("Eg_in_R"), c.arr("Eg_in_T"), c.arr("Cb_L"),
                       c.arr("Cb_B"), c.arr("Cb_R"), c.arr("Cb_T"));
    TRT_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// GREY_COEFS, GLOQD_SOLVES.f08:89-143
// ---------------------------------------------------------------------------
struct GCArgs {
    int nx, ny, ng;
    double c, dt, th, Fdt;
    const double *A, *Ea, *Ev, *Eh, *Fxo, *Fyo, *fxx, *fyy, *fvxx, *fvxy, *fhyy, *fhxy, *KE, *KR, *PoL, *PoR, *PoB, *PoT;
    double *KB, *DCxx, *DLxx, *DRxx, *DCyy, *DByy, *DTyy, *DLxy, *DRxy, *DBxy, *DTxy, *PL, *PR, *PB, *PT;
};

__global__ void grey_coef_kernel(GCArgs a)
{
    const size_t nc = (size_t)a.nx * a.ny, nv = (size_t)(a.nx + 1) * a.ny, nh = (size_t)a.nx * (a.ny + 1);
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < nc; k += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(k % a.nx), j = (int)(k / a.nx);
        const size_t v0 = i + (size_t)(a.nx + 1) * j, h0 = i + (size_t)a.nx * j;
        const double A = a.A[k];
        double kb = 0, dcxx = 0, dlxx = 0, drxx = 0, dcyy = 0, dbyy = 0, dtyy = 0, dlxy = 0, drxy = 0, dbxy = 0, dtxy = 0;
        double pl = 0, pr = 0, pb = 0, pt = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0;
        for (int g = 0; g < a.ng; ++g) {
            const size_t kg = k + g * nc, vL = v0 + g * nv, vR = vL + 1, hB = h0 + g * nh, hT = hB + a.nx;
            const double tk = a.KR[kg] + a.Fdt / (a.th * a.c * a.dt);
            kb += a.KE[kg] * a.Ea[kg];
            dcxx += a.fxx[kg] * a.Ea[kg] / tk;
            dlxx += a.fvxx[vL] * a.Ev[vL] / tk;
            drxx += a.fvxx[vR] * a.Ev[vR] / tk;
            dcyy += a.fyy[kg] * a.Ea[kg] / tk;
            dbyy += a.fhyy[hB] * a.Eh[hB] / tk;
            dtyy += a.fhyy[hT] * a.Eh[hT] / tk;
            dlxy += a.fvxy[vL] * a.Ev[vL] / tk;
            drxy += a.fvxy[vR] * a.Ev[vR] / tk;
            dbxy += a.fhxy[hB] * a.Eh[hB] / tk;
            dtxy += a.fhxy[hT] * a.Eh[hT] / tk;
            const double w = a.Fdt * A / (2.0 * a.th * a.c * a.dt);
            pl += (w * a.Fxo[vL] + a.PoL[kg]) / tk;
            pr += (w * a.Fxo[vR] + a.PoR[kg]) / tk;
            pb += (w * a.Fyo[hB] + a.PoB[kg]) / tk;
            pt += (w * a.Fyo[hT] + a.PoT[kg]) / tk;
            s1 += a.Ea[kg]; s2 += a.Ev[vL]; s3 += a.Ev[vR]; s4 += a.Eh[hB]; s5 += a.Eh[hT];
        }
        a.KB[k] = kb / s1;
        a.DCxx[k] = dcxx / s1; a.DLxx[k] = dlxx / s2; a.DRxx[k] = drxx / s3;
        a.DCyy[k] = dcyy / s1; a.DByy[k] = dbyy / s4; a.DTyy[k] = dtyy / s5;
        a.DLxy[k] = dlxy / s2; a.DRxy[k] = drxy / s3; a.DBxy[k] = dbxy / s4; a.DTxy[k] = dtxy / s5;
        a.PL[k] = pl; a.PR[k] = pr; a.PB[k] = pb; a.PT[k] = pt;
    }
}

void grey_coefs(Ctx &c)
{
    GCArgs a{};
    a.nx = c.nx; a.ny = c.ny; a.ng = c.ng;
    a.c = c.p.c; a.dt = c.p.dt; a.th = c.p.theta; a.Fdt = 1.0;
    a.A = c.A;
    a.Ea = c.arr("Eg_avg"); a.Ev = c.arr("Eg_edgV"); a.Eh = c.arr("Eg_edgH");
    a.Fxo = c.arr("Fxg_edgV_old"); a.Fyo = c.arr("Fyg_edgH_old");
    a.fxx = c.arr("fg_avg_xx"); a.fyy = c.arr("fg_avg_yy"); a.fvxx = c.arr("fg_edgV_xx"); a.fvxy = c.arr("fg_edgV_xy");
    a.fhyy = c.arr("fg_edgH_yy"); a.fhxy = c.arr("fg_edgH_xy");
    a.KE = c.arr("KapE"); a.KR = c.arr("KapR");
    a.PoL = c.arr("Pold_L"); a.PoR = c.arr("Pold_R"); a.PoB = c.arr("Pold_B"); a.PoT = c.arr("Pold_T");
    a.KB = c.arr("KapE_Bar");
    a.DCxx = c.arr("DC_xx"); a.DLxx = c.arr("DL_xx"); a.DRxx = c.arr("DR_xx");
    a.DCyy = c.arr("DC_yy"); a.DByy = c.arr("DB_yy"); a.DTyy = c.arr("DT_yy");
    a.DLxy = c.arr("DL_xy"); a.DRxy = c.arr("DR_xy"); a.DBxy = c.arr("DB_xy"); a.DTxy = c.arr("DT_xy");
    a.PL = c.arr("PL"); a.PR = c.arr("PR"); a.PB = c.arr("PB"); a.PT = c.arr("PT");
    hipLaunchKernelGGL(grey_coef_kernel, dim3(nb((size_t)c.nx * c.ny)), dim3(256), 0, c.stream, a);
    TRT_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// EGP_FV_NEWT, GLOQD_SOLVES.f08:305-860
// ---------------------------------------------------------------------------
struct NewtArgs {
    int nx, ny, nbd, damp;
    double c, th, dt, cv, chi, qfac, Fin, Ein;
    // grey_emission 1: Q_bar from the group sources (opacity model, groups, multipliers)
    int gem, ng, opm;
    double cu;
    const double *nu, *kmult;
    const int *mat;
    const double *A, *dx, *dy, *Rh, *Gh;
    const double *DCxx, *DLxx, *DRxx, *DCyy, *DByy, *DTyy, *DLxy, *DRxy, *DBxy, *DTxy, *PL, *PR, *PB, *PT;
    const double *CbL, *CbB, *CbR, *CbT, *EinL, *EinB, *EinR, *EinT, *FinL, *FinB, *FinR, *FinT;
    double *T, *Ea, *Ev, *Eh, *Fx, *Fy, *KB, *KBdT, *KdTorg, *Qb;
    double *drT, *drB, *drML, *drMR, *drMB, *drMT, *W;
    double *dT, *dEa, *dEv, *dEh, *dFx, *dFy;
    double *coef, *bnd;
};

// grey_emission 1 (a labelled deviation, off by default): the grey emission consistent
// with MATERIAL_UPDATE's group sources (UPDATES.f08:130-152), Q_bar(T) = sum_g 4 pi
// kappa_g(T) mult B_g(T) in cell k, and its T-derivative by a central difference of step
// 1e-4 T (the derivative steers the Newton iteration; its fixed point does not depend on it)
__device__ double qbar_groups(const NewtArgs &a, size_t k, double T)
{
    const double mult = a.kmult[a.mat[k]];
    double q = 0.0;
    for (int g = 0; g < a.ng; ++g) {
        const double nl = a.nu[g], nh = a.nu[g + 1];
        const double kap = a.opm == 0 ? d_kapB(T, nl, nh) : d_kap_mid(T, nl, nh);
        q += 4.0 * PI * (kap * mult) * d_bg_planck(T, nl, nh, a.cu);
    }
    return q;
}

__device__ void qbar(const NewtArgs &a, size_t k, double T, double *Qb, double *QbdT)
{
    if (a.gem == 1) {
        const double h = 1e-4 * T;
        *Qb = qbar_groups(a, k, T);
        *QbdT = (qbar_groups(a, k, T + h) - qbar_groups(a, k, T - h)) / (2.0 * h);
    } else {
        *Qb = a.qfac * T;        // GL:439-440
        *QbdT = a.qfac;
    }
}

// residuals and cell coefficients (GL:437-513)
__global__ void newt_cells(NewtArgs a)
{
    const size_t nc = (size_t)a.nx * a.ny;
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < nc; k += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(k % a.nx), j = (int)(k / a.nx);
        const size_t vL = i + (size_t)(a.nx + 1) * j, vR = vL + 1, hB = i + (size_t)a.nx * j, hT = hB + a.nx;
        const double A = a.A[k], dx = a.dx[i], dy = a.dy[j], c = a.c, th = a.th, dt = a.dt, cv = a.cv;
        const double T = a.T[k], Ea = a.Ea[k], KB = a.KB[k];
        double Qb, QbdT;
        qbar(a, k, T, &Qb, &QbdT);
        a.Qb[k] = Qb;
        double kdt = a.KBdT[k];
        if (a.damp) {                                   // GL:444-454
            kdt = a.KdTorg[k];
            const double bound = (cv / (th * dt) + QbdT) / (c * Ea);
            if (kdt > a.chi * bound) kdt = a.chi * bound;
            a.KBdT[k] = kdt;
        }
        const double drT = cv / (th * dt) * T + Qb - c * KB * Ea - a.Rh[k];
        const double drB = A * (1.0 / (th * dt) + c * KB) * Ea + dy * (a.Fx[vR] - a.Fx[vL]) + dx * (a.Fy[hT] - a.Fy[hB])
                           - A * Qb - a.Gh[k];
        const double drML = A * a.Fx[vL] / 2.0 + c * dy * (a.DCxx[k] * Ea - a.DLxx[k] * a.Ev[vL])
                            + c * dx * (a.DTxy[k] * a.Eh[hT] - a.DBxy[k] * a.Eh[hB]) / 2.0 - a.PL[k];
        const double drMR = A * a.Fx[vR] / 2.0 + c * dy * (a.DRxx[k] * a.Ev[vR] - a.DCxx[k] * Ea)
                            + c * dx * (a.DTxy[k] * a.Eh[hT] - a.DBxy[k] * a.Eh[hB]) / 2.0 - a.PR[k];
        const double drMB = A * a.Fy[hB] / 2.0 + c * dx * (a.DCyy[k] * Ea - a.DByy[k] * a.Eh[hB])
                            + c * dy * (a.DRxy[k] * a.Ev[vR] - a.DLxy[k] * a.Ev[vL]) / 2.0 - a.PB[k];
        const double drMT = A * a.Fy[hT] / 2.0 + c * dx * (a.DTyy[k] * a.Eh[hT] - a.DCyy[k] * Ea)
                            + c * dy * (a.DRxy[k] * a.Ev[vR] - a.DLxy[k] * a.Ev[vL]) / 2.0 - a.PT[k];
        const double W = cv / (th * dt) + QbdT - c * kdt * Ea;
        const double drG = -drB + drT * (A / W) * (c * kdt * Ea - QbdT) + 2.0 * dy * (drMR - drML) / A
                           + 2.0 * dx * (drMT - drMB) / A;
        a.drT[k] = drT; a.drB[k] = drB; a.drML[k] = drML; a.drMR[k] = drMR; a.drMB[k] = drMB; a.drMT[k] = drMT;
        a.W[k] = W;
        double *cf = a.coef;
        cf[EBC * nc + k] = A * (1.0 / (th * dt) + c * KB * (1.0 + (c * kdt * Ea - QbdT) / W))
                           + 4.0 * c * dy * dy * a.DCxx[k] / A + 4.0 * c * dx * dx * a.DCyy[k] / A;
        cf[EBL * nc + k] = -2.0 * c * dy * dy * a.DLxx[k] / A;
        cf[EBR * nc + k] = -2.0 * c * dy * dy * a.DRxx[k] / A;
        cf[EBB * nc + k] = -2.0 * c * dx * dx * a.DByy[k] / A;
        cf[EBT * nc + k] = -2.0 * c * dx * dx * a.DTyy[k] / A;
        cf[MXC * nc + k] = 2.0 * c * dy * a.DCxx[k] / A;
        cf[MXT * nc + k] = -c * dx * a.DTxy[k] / A;
        cf[MXB * nc + k] = c * dx * a.DBxy[k] / A;
        cf[MYC * nc + k] = 2.0 * c * dx * a.DCyy[k] / A;
        cf[MYR * nc + k] = -c * dy * a.DRxy[k] / A;
        cf[MYL * nc + k] = c * dy * a.DLxy[k] / A;
        cf[ERHS * nc + k] = drG;
    }
}

// face coefficients and boundary rows (GL:515-557)
__global__ void newt_faces(NewtArgs a)
{
    const size_t nc = (size_t)a.nx * a.ny;
    const double c = a.c;
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < nc; k += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(k % a.nx), j = (int)(k / a.nx);
        double *cf = a.coef;
        if (j < a.ny - 1) {
            const size_t ku = k + a.nx;
            cf[MYT * nc + k] = -2.0 * c * a.dx[i] * (a.DTyy[k] / a.A[k] + a.DByy[ku] / a.A[ku]);
            cf[MYRHS * nc + k] = 2.0 * a.drMT[k] / a.A[k] - 2.0 * a.drMB[ku] / a.A[ku];
        }
        if (i < a.nx - 1) {
            const size_t kr = k + 1;
            cf[MXR * nc + k] = -2.0 * c * a.dy[j] * (a.DRxx[k] / a.A[k] + a.DLxx[kr] / a.A[kr]);
            cf[MXRHS * nc + k] = 2.0 * a.drMR[k] / a.A[k] - 2.0 * a.drML[kr] / a.A[kr];
        }
        double *bd = a.bnd;
        const int nbd = a.nbd;
        if (i == 0) {
            bd[CPL * nbd + j] = c * a.CbL[j] - 2.0 * c * a.dy[j] * a.DLxx[k] / a.A[k];
            bd[BCL * nbd + j] = c * a.CbL[j] * (a.Ein * a.EinL[j] - a.Ev[(size_t)(a.nx + 1) * j]) + a.Fx[(size_t)(a.nx + 1) * j]
                                - a.Fin * a.FinL[j] - 2.0 * a.drML[k] / a.A[k];
        }
        if (i == a.nx - 1) {
            const size_t v = a.nx + (size_t)(a.nx + 1) * j;
            bd[CPR * nbd + j] = c * a.CbR[j] + 2.0 * c * a.dy[j] * a.DRxx[k] / a.A[k];
            bd[BCR * nbd + j] = c * a.CbR[j] * (a.Ein * a.EinR[j] - a.Ev[v]) + a.Fx[v] - a.Fin * a.FinR[j]
                                - 2.0 * a.drMR[k] / a.A[k];
        }
        if (j == 0) {
            bd[CPB * nbd + i] = c * a.CbB[i] - 2.0 * c * a.dx[i] * a.DByy[k] / a.A[k];
            bd[BCB * nbd + i] = c * a.CbB[i] * (a.Ein * a.EinB[i] - a.Eh[i]) + a.Fy[i] - a.Fin * a.FinB[i]
                                - 2.0 * a.drMB[k] / a.A[k];
        }
        if (j == a.ny - 1) {
            const size_t h = i + (size_t)a.nx * a.ny;
            bd[CPT * nbd + i] = c * a.CbT[i] + 2.0 * c * a.dx[i] * a.DTyy[k] / a.A[k];
            bd[BCT * nbd + i] = c * a.CbT[i] * (a.Ein * a.EinT[i] - a.Eh[h]) + a.Fy[h] - a.Fin * a.FinT[i]
                                - 2.0 * a.drMT[k] / a.A[k];
        }
    }
}

// Delta T and Delta F from Delta E (GL:573-597)
__global__ void newt_deltas(NewtArgs a)
{
    const size_t nc = (size_t)a.nx * a.ny;
    const double c = a.c;
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < nc; k += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(k % a.nx), j = (int)(k / a.nx);
        const size_t vL = i + (size_t)(a.nx + 1) * j, vR = vL + 1, hB = i + (size_t)a.nx * j, hT = hB + a.nx;
        const double A = a.A[k], dx = a.dx[i], dy = a.dy[j];
        a.dT[k] = (c * a.KB[k] * a.dEa[k] - a.drT[k]) / a.W[k];
        // the reference uses DT_xy for both horizontal terms here (GL:579-580, 588-589)
        a.dFx[vL] = -2.0 * c * dy * (a.DCxx[k] * a.dEa[k] - a.DLxx[k] * a.dEv[vL]) / A
                    - c * dx * (a.DTxy[k] * a.dEh[hT] - a.DTxy[k] * a.dEh[hB]) / A - 2.0 * a.drML[k] / A;
        a.dFy[hB] = -2.0 * c * dx * (a.DCyy[k] * a.dEa[k] - a.DByy[k] * a.dEh[hB]) / A
                    - c * dy * (a.DRxy[k] * a.dEv[vR] - a.DLxy[k] * a.dEv[vL]) / A - 2.0 * a.drMB[k] / A;
        if (i == a.nx - 1)
            a.dFx[vR] = -2.0 * c * dy * (a.DRxx[k] * a.dEv[vR] - a.DCxx[k] * a.dEa[k]) / A
                        - c * dx * (a.DTxy[k] * a.dEh[hT] - a.DTxy[k] * a.dEh[hB]) / A - 2.0 * a.drMR[k] / A;
        if (j == a.ny - 1)
            a.dFy[hT] = -2.0 * c * dx * (a.DTyy[k] * a.dEh[hT] - a.DCyy[k] * a.dEa[k]) / A
                        - c * dy * (a.DRxy[k] * a.dEv[vR] - a.DLxy[k] * a.dEv[vL]) / A - 2.0 * a.drMT[k] / A;
    }
}

// sum of squares of several arrays (line-search norms), deterministic two-level reduction
__global__ void sumsq_kernel(const double *x, size_t n, double *part)
{
    __shared__ double sh[8];
    double s = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) s += x[i] * x[i];
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void maxabs_kernel2(const double *x, size_t n, double *part)
{
    __shared__ double sh[256];
    double m = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) m = fmax(m, fabs(x[i]));
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void scale_kernel(double *x, size_t n, double s)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) x[i] /= s;
}

__global__ void axpy1_kernel(double *y, const double *x, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) y[i] += x[i];
}

// T, E update, KapE_Bar += dKapE_Bar/dT * dT, Q_bar (GL:759-766)
__global__ void newt_update(NewtArgs a)
{
    const size_t nc = (size_t)a.nx * a.ny;
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < nc; k += (size_t)gridDim.x * blockDim.x) {
        a.T[k] += a.dT[k];
        a.Ea[k] += a.dEa[k];
        a.KB[k] = a.KB[k] + a.KBdT[k] * a.dT[k];
        a.Qb[k] = a.gem == 1 ? qbar_groups(a, k, a.T[k]) : a.qfac * a.T[k];   // GL:766
    }
}

// Positivity safety search (GL:643-752): per checked value x with step d, the number of
// times the step length must be divided by line_src before x + d / LS reaches the lower
// bound, the step length built by repeated multiplication as the reference builds it.
// The reference divides every delta by each cell's LS in turn, so the total division is
// line_src^(largest count); counts are combined by atomicMax into *emax.  A value already
// below its bound never reaches it (the reference loops forever): the count is capped at
// kSafeCap and reported.
constexpr int kSafeCap = 4096;
__global__ void safety_count_kernel(const double *x, const double *d, size_t n, double bound, double line_src,
                                    int *emax)
{
    int best = 0;
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
        double LS = 1.0;
        int p = 0;
        while (x[k] + d[k] / LS < bound && p < kSafeCap) {
            LS *= line_src;
            ++p;
        }
        best = max(best, p);
    }
    if (best > 0) atomicMax(emax, best);
}

// dKapE_Bar/dT by finite difference over MLOQD iterates (GL:402-417)
__global__ void kbdt_kernel(size_t n, const double *T, const double *Tm, const double *KB, const double *KBm, double *out,
                            double *org)
{
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
        const double v = fabs(T[k] - Tm[k]) < 1e-15 ? 0.0 : (KB[k] - KBm[k]) / (T[k] - Tm[k]);
        out[k] = v;
        org[k] = v;
    }
}

int egp_fv_newt(Ctx &c, int mgqd_it)
{
    if (c.grey.nsys == 0) qd_alloc(c, c.grey, 1);
    c.newt_kits.clear();
    const int nx = c.nx, ny = c.ny;
    const size_t nc = (size_t)nx * ny, nv = (size_t)(nx + 1) * ny, nh = (size_t)nx * (ny + 1);
    auto get = [&](const char *n, int d0, int d1) -> double * {
        auto it = c.arrays.find(n);
        if (it != c.arrays.end()) return it->second.p;
        return c.reg(n, d0, d1).p;
    };
    // sig_R and aR as GL:371-372 (hard-coded h, c, erg); Q_bar = 15 kap0 c aR T / pi^4
    const double cc = C_CONST, h = H_PLANCK;
    const double sig_R = 2.0 * std::pow(PI, 5) / (15.0 * cc * cc * h * h * h * ERG * ERG * ERG * ERG * c.p.comp_unit);
    const double aR = 4.0 * sig_R / cc;
    NewtArgs a{};
    a.nx = nx; a.ny = ny; a.nbd = std::max(nx, ny);
    a.c = cc; a.th = c.p.theta; a.dt = c.p.dt; a.cv = c.p.cv; a.chi = c.p.chi;
    a.qfac = 15.0 * c.p.kap0 * cc * aR / (PI * PI * PI * PI);
    a.Fin = 1.0; a.Ein = 1.0;
    a.gem = c.p.grey_emission; a.ng = c.ng; a.opm = c.p.opacity_model; a.cu = c.p.comp_unit;
    a.nu = c.nu; a.kmult = c.kmult; a.mat = c.mat;
    a.A = c.A; a.dx = c.dx; a.dy = c.dy; a.Rh = c.arr("Rhat_old"); a.Gh = c.arr("Gold_Hat");
    a.DCxx = c.arr("DC_xx"); a.DLxx = c.arr("DL_xx"); a.DRxx = c.arr("DR_xx");
    a.DCyy = c.arr("DC_yy"); a.DByy = c.arr("DB_yy"); a.DTyy = c.arr("DT_yy");
    a.DLxy = c.arr("DL_xy"); a.DRxy = c.arr("DR_xy"); a.DBxy = c.arr("DB_xy"); a.DTxy = c.arr("DT_xy");
    a.PL = c.arr("PL"); a.PR = c.arr("PR"); a.PB = c.arr("PB"); a.PT = c.arr("PT");
    a.CbL = c.arr("Cb_L"); a.CbB = c.arr("Cb_B"); a.CbR = c.arr("Cb_R"); a.CbT = c.arr("Cb_T");
    a.EinL = c.arr("E_in_L"); a.EinB = c.arr("E_in_B"); a.EinR = c.arr("E_in_R"); a.EinT = c.arr("E_in_T");
    a.FinL = c.arr("F_in_L"); a.FinB = c.arr("F_in_B"); a.FinR = c.arr("F_in_R"); a.FinT = c.arr("F_in_T");
    a.T = c.arr("Temp"); a.Ea = c.arr("E_avg"); a.Ev = c.arr("E_edgV"); a.Eh = c.arr("E_edgH");
    a.Fx = c.arr("Fx_edgV"); a.Fy = c.arr("Fy_edgH"); a.KB = c.arr("KapE_Bar"); a.KBdT = c.arr("KapE_Bar_dT");
    a.Qb = c.arr("GQD_Src");
    a.KdTorg = get("_kdT_org", nx, ny);
    a.drT = get("dr_T", nx, ny); a.drB = get("dr_B", nx, ny); a.drML = get("dr_ML", nx, ny);
    a.drMR = get("dr_MR", nx, ny); a.drMB = get("dr_MB", nx, ny); a.drMT = get("dr_MT", nx, ny);
    a.W = get("_W", nx, ny);
    a.dT = get("Del_T", nx, ny); a.dEa = get("Del_E_avg", nx, ny); a.dEv = get("Del_E_edgV", nx + 1, ny);
    a.dEh = get("Del_E_edgH", nx, ny + 1); a.dFx = get("Del_Fx_edgV", nx + 1, ny); a.dFy = get("Del_Fy_edgH", nx, ny + 1);
    double *old[6] = {get("_dT_o", nx, ny), get("_dEa_o", nx, ny), get("_dEv_o", nx + 1, ny),
                      get("_dEh_o", nx, ny + 1), get("_dFx_o", nx + 1, ny), get("_dFy_o", nx, ny + 1)};
    double *del[6] = {a.dT, a.dEa, a.dEv, a.dEh, a.dFx, a.dFy};
    const size_t sz[6] = {nc, nc, nv, nh, nv, nh};
    double *val[6] = {a.T, a.Ea, a.Ev, a.Eh, a.Fx, a.Fy};
    a.coef = c.grey.coef; a.bnd = c.grey.bnd;
    // deltas start at zero every call (GL:394-399)
    for (int q = 0; q < 6; ++q) dfill(c, del[q], 0.0, sz[q]);
    const bool dTflag = mgqd_it > 1 && c.p.kapE_dT_flag;
    if (dTflag)
        hipLaunchKernelGGL(kbdt_kernel, dim3(nb(nc)), dim3(256), 0, c.stream, nc, a.T, c.arr("Temp_MGQDold2"), a.KB,
                           c.arr("KapE_bar_MGQDold"), a.KBdT, a.KdTorg);
    else
        dfill(c, a.KBdT, 0.0, nc);
    a.damp = dTflag ? 1 : 0;
    const int PB = 64;   // reduction blocks
    double *part = c.scratch;        // [12][PB]
    int its = 0;
    bool converged = false;
    while (!converged && its < c.p.maxit_gloqd) {
        its++;
        for (int q = 0; q < 6; ++q) dcopy(c, old[q], del[q], sz[q]);
        hipLaunchKernelGGL(newt_cells, dim3(nb(nc)), dim3(256), 0, c.stream, a);
        hipLaunchKernelGGL(newt_faces, dim3(nb(nc)), dim3(256), 0, c.stream, a);
        // previous Delta E is the initial guess (GL:565)
        hipLaunchKernelGGL(qd_gather, dim3(nb(c.grey.nrow)), dim3(256), 0, c.stream, c.grey.x, a.dEa, a.dEv, a.dEh, nc, nv, nh, 1);
        // the deltas are measured against the grey E they correct
        hipLaunchKernelGGL(qd_gather, dim3(nb(c.grey.nrow)), dim3(256), 0, c.stream, c.grey.ref, a.Ea, a.Ev, a.Eh, nc, nv, nh, 1);
        TRT_CHECK(hipGetLastError());
        int kits = 0;
        qd_solve(c, c.grey, &kits, true);
        c.newt_kits.push_back(kits);
        hipLaunchKernelGGL(qd_scatter, dim3(nb(c.grey.nrow)), dim3(256), 0, c.stream, c.grey.x, a.dEa, a.dEv, a.dEh, nc, nv, nh, 1);
        hipLaunchKernelGGL(newt_deltas, dim3(nb(nc)), dim3(256), 0, c.stream, a);
        // LINE_SEARCH on the six delta norms in the reference order (GL:604-636)
        if (c.p.use_line_search && its > 1) {
            for (int q = 0; q < 6; ++q) {
                hipLaunchKernelGGL(sumsq_kernel, dim3(PB), dim3(256), 0, c.stream, del[q], sz[q], part + (2 * q) * PB);
                hipLaunchKernelGGL(sumsq_kernel, dim3(PB), dim3(256), 0, c.stream, old[q], sz[q], part + (2 * q + 1) * PB);
            }
            TRT_CHECK(hipMemcpyAsync(c.hbuf, part, 12 * PB * 8, hipMemcpyDeviceToHost, c.stream));
            TRT_CHECK(hipStreamSynchronize(c.stream));
            double S = 1.0;
            for (int q = 0; q < 6; ++q) {
                double n1 = 0, n0 = 0;
                for (int k = 0; k < PB; ++k) { n1 += c.hbuf[(2 * q) * PB + k]; n0 += c.hbuf[(2 * q + 1) * PB + k]; }
                const double norm = std::sqrt(n1) / S, norm_old = std::sqrt(n0);
                double LS = 1.0;
                for (int t = 0; t < 10; ++t) {
                    if (norm / LS > norm_old) { LS *= c.p.line_src; continue; }
                    break;
                }
                if (LS > 1.0) S *= LS;
            }
            if (S > 1.0)
                for (int q = 0; q < 6; ++q)
                    hipLaunchKernelGGL(scale_kernel, dim3(nb(sz[q])), dim3(256), 0, c.stream, del[q], sz[q], S);
        }
        if (c.p.use_safety_search) {
            // T and E_avg (cells), E_edgV, E_edgH against their lower bounds (GL:643-752)
            int *emax = reinterpret_cast<int *>(part + 12 * PB);
            TRT_CHECK(hipMemsetAsync(emax, 0, sizeof(int), c.stream));
            const double *xs[4] = {a.T, a.Ea, a.Ev, a.Eh};
            const double *ds[4] = {a.dT, a.dEa, a.dEv, a.dEh};
            const size_t ns[4] = {nc, nc, nv, nh};
            const double bnd[4] = {c.p.T_bound_low, c.p.E_bound_low, c.p.E_bound_low, c.p.E_bound_low};
            for (int q = 0; q < 4; ++q)
                hipLaunchKernelGGL(safety_count_kernel, dim3(std::min(nb(ns[q]), 1024)), dim3(256), 0, c.stream, xs[q],
                                   ds[q], ns[q], bnd[q], c.p.line_src, emax);
            int e = 0;
            TRT_CHECK(hipMemcpyAsync(&e, emax, sizeof(int), hipMemcpyDeviceToHost, c.stream));
            TRT_CHECK(hipStreamSynchronize(c.stream));
            if (e >= kSafeCap)
                throw Error("EGP_FV_NEWT safety search: a temperature or energy density is already below its lower "
                            "bound, so no step length reaches it");
            if (e > 0) {
                double S = 1.0;
                for (int t = 0; t < e; ++t) S *= c.p.line_src;
                for (int q = 0; q < 6; ++q)
                    hipLaunchKernelGGL(scale_kernel, dim3(nb(sz[q])), dim3(256), 0, c.stream, del[q], sz[q], S);
            }
        }
        hipLaunchKernelGGL(newt_update, dim3(nb(nc)), dim3(256), 0, c.stream, a);
        for (int q = 2; q < 6; ++q)
            hipLaunchKernelGGL(axpy1_kernel, dim3(nb(sz[q])), dim3(256), 0, c.stream, val[q], del[q], sz[q]);
        // convergence (GL:773-793)
        for (int q = 0; q < 6; ++q) {
            hipLaunchKernelGGL(maxabs_kernel2, dim3(PB), dim3(256), 0, c.stream, del[q], sz[q], part + (2 * q) * PB);
            hipLaunchKernelGGL(maxabs_kernel2, dim3(PB), dim3(256), 0, c.stream, val[q], sz[q], part + (2 * q + 1) * PB);
        }
        TRT_CHECK(hipMemcpyAsync(c.hbuf, part, 12 * PB * 8, hipMemcpyDeviceToHost, c.stream));
        TRT_CHECK(hipStreamSynchronize(c.stream));
        converged = true;
        for (int q = 0; q < 6; ++q) {
            double md = 0, mv = 0;
            for (int k = 0; k < PB; ++k) { md = std::max(md, c.hbuf[(2 * q) * PB + k]); mv = std::max(mv, c.hbuf[(2 * q + 1) * PB + k]); }
            if (md > c.p.conv_gr1 * mv + c.p.conv_gr2) converged = false;
        }
    }
    return its;
}

// ---------------------------------------------------------------------------
// MEB_SOLVE, UPDATES.f08:189-231: closed-form grey temperature (use_grey = 0)
// ---------------------------------------------------------------------------
__global__ void meb_kernel(double *T, const double *Eg, const double *KapE, const double *Told, size_t n, int ng,
                           double fac, double dt, double cv)
{
    const double cc = C_CONST;
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
        double E = 0.0, kb = 0.0;
        for (int g = 0; g < ng; ++g) {
            E += Eg[k + g * n];
            kb += KapE[k + g * n] * Eg[k + g * n];
        }
        kb /= E;
        T[k] = fac * (dt * cc * kb * E + cv * Told[k]);
    }
}

void meb_solve(Ctx &c, const char *eg)
{
    const size_t n = (size_t)c.nx * c.ny;
    auto it = c.arrays.find(eg);
    if (it == c.arrays.end() || it->second.size() != n * c.ng)
        throw Error(std::string("MEB_SOLVE needs a group energy array (N_x,N_y,N_g), not ") + eg);
    // (c^2 h^3)/(c^2 h^3 cv cu + 8 pi dt kap0 / erg^4) ... * cu, with the hard-coded c (UP:196)
    const double cc = C_CONST, h3 = H_PLANCK * H_PLANCK * H_PLANCK, e4 = ERG * ERG * ERG * ERG;
    const double fac = (cc * cc * h3) / (cc * cc * h3 * c.p.cv * c.p.comp_unit + 8.0 * PI * c.p.dt * c.p.kap0 / e4)
                       * c.p.comp_unit;
    hipLaunchKernelGGL(meb_kernel, dim3(nb(n)), dim3(256), 0, c.stream, c.arr("Temp"), it->second.p, c.arr("KapE"),
                       c.arr("Temp_Old"), n, c.ng, fac, c.p.dt, c.p.cv);
    TRT_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// one time step, TRT_MLQD_ALGORITHM (mlqd), ALGORITHMS.f08:389-897
// ---------------------------------------------------------------------------
static void cp(Ctx &c, const char *dst, const char *src)
{
    auto &d = c.arrays.at(dst);
    dcopy(c, d.p, c.arr(src), d.size());
}

// The time step of TRT_MLQD_ALGORITHM in phases, so that a batch can run the sweeps of
// many problems' RT iterations as one launch: step_begin (the step's prologue), then
// step_rt_iteration until step_done, then step_end.
void step_begin(Ctx &c, StepState &st, int ho_form)
{
    const size_t nc = (size_t)c.nx * c.ny;
    st = StepState{};
    st.ho_form = ho_form;
    c.time += c.p.dt;
    // AL:396-414 (I_crn_old = I_crn is a buffer handover)
    cp(c, "Temp_Old", "Temp");
    c.iold = c.icur;
    cp(c, "fg_avg_xx_old", "fg_avg_xx"); cp(c, "fg_avg_yy_old", "fg_avg_yy");
    cp(c, "fg_edgV_xx_old", "fg_edgV_xx"); cp(c, "fg_edgV_xy_old", "fg_edgV_xy");
    cp(c, "fg_edgH_yy_old", "fg_edgH_yy"); cp(c, "fg_edgH_xy_old", "fg_edgH_xy");
    cp(c, "MGQD_Src_old", "MGQD_Src"); cp(c, "GQD_Src_old", "GQD_Src");
    cp(c, "E_avg_old", "E_avg"); cp(c, "Eg_avg_old", "Eg_avg");
    cp(c, "Eg_edgV_old", "Eg_edgV"); cp(c, "Eg_edgH_old", "Eg_edgH");
    cp(c, "Fxg_edgV_old", "Fxg_edgV"); cp(c, "Fyg_edgH_old", "Fyg_edgH");
    cp(c, "KapE_Bar_old", "KapE_Bar"); cp(c, "KapE_old", "KapE"); cp(c, "KapR_old", "KapR");
    old_mgqd_coefs(c);                       // AL:417-419
    old_grey_coefs(c);                       // AL:420
    for (auto n : {"HO_E_avg_RTold2", "HO_E_avg_RTold", "Temp_RTold2", "Temp_RTold"}) dfill(c, c.arr(n), 0.0, nc);
    c.itlog.clear();
    // factors are reused within a time step only: a step depends on the state at its
    // start alone, so a run resumed from a restart file continues bitwise
    qd_nd_forget(c.mg);
    qd_nd_forget(c.grey);
    // ho_form 1 (the ROM runs, HO_Form 1 and 3, AL:573-631): in each RT iteration the
    // QD factors and C_g of the new time level come from the caller's reconstruction
    // (pushed to the *_next arrays) where POD/DMD_RECONSTRUCT_fg would write them; one RT
    // iteration, as INPUT sets maxit_RTE = 1 for 'fg' ROM runs (IP:94-104)
    st.maxit_rte = ho_form ? 1 : c.p.maxit_rte;
}

bool step_done(const StepState &st) { return st.rt_conv || st.rt_its >= st.maxit_rte; }

// the next RT iteration starts with a sweep (from the second on, AL:302)
bool step_sweeps_next(const StepState &st) { return !step_done(st) && !st.ho_form && st.rt_its + 1 >= 2; }

// One RT iteration (AL:527-837).  swept: its sweep has been enqueued already (by a
// batched sweep that the context's stream waits for).
void step_rt_iteration(Ctx &c, StepState &st, bool swept)
{
    const size_t nc = (size_t)c.nx * c.ny;
    const bool sweeps = step_sweeps_next(st);
    st.rt_its++;
    if (sweeps) {
        if (!swept) transport_sweep(c);      // AL:541-542
        st.sweeps++;
        collapse_with_fg(c);                 // AL:551-553 and fg_Calc AL:568-569
        if (c.refl) {                        // AL:556-565 (the moments above already hold the
            rt_bc_update(c);                 // sweep's inflow; the update feeds the next sweep)
            mgqd_in_calc(c);
            if (!c.p.no_grey) gqd_in_calc(c);
        }
        cg_calc(c);                          // AL:572
    } else if (st.ho_form) {
        for (auto n : {"fg_avg_xx", "fg_avg_yy", "fg_edgV_xx", "fg_edgV_xy", "fg_edgH_yy", "fg_edgH_xy", "Cg_L", "Cg_B",
                       "Cg_R", "Cg_T"})
            cp(c, n, (std::string(n) + "_next").c_str());   // AL:582-593
    }
    dfill(c, c.arr("MGQD_E_avg_MGQDold2"), 0.0, nc);   // AL:676-679
    cp(c, "MGQD_E_avg_MGQDold", "HO_E_avg");
    dfill(c, c.arr("Temp_MGQDold2"), 0.0, nc);
    cp(c, "Temp_MGQDold", "Temp");
    int mg_its = 0;
    bool mg_conv = false;
    c.itlog.emplace_back();
    RtItLog &rlog = c.itlog.back();
    while (!mg_conv && mg_its < c.p.maxit_mloqd) {
        mg_its++;
        rlog.mgl.emplace_back();
        MgItLog &mlog = rlog.mgl.back();
        mloqd_fv(c, nullptr);                // AL:686-689
        collapse_mg_ef(c);                   // AL:692-693
        if (c.p.no_grey) {
            meb_solve(c, "Eg_avg");          // AL:767-770
        } else {
            cbar_calc(c);                    // AL:709
            cp(c, "KapE_bar_MGQDold", "KapE_Bar");   // AL:712
            grey_coefs(c);                   // AL:713-715
            cp(c, "E_avg", "MGQD_E_avg"); cp(c, "E_edgV", "MGQD_E_edgV"); cp(c, "E_edgH", "MGQD_E_edgH");
            cp(c, "Fx_edgV", "MGQD_Fx_edgV"); cp(c, "Fy_edgH", "MGQD_Fy_edgH");   // AL:717-721
            mlog.newt = egp_fv_newt(c, mg_its);      // AL:724-728
            mlog.gkits = c.newt_kits;
            st.newt_total += mlog.newt;
        }
        material_update(c);                  // AL:774
        int tc = 0, ec = 0;
        double tn, tr, en, er;
        convergence(c, c.arr("Temp"), c.arr("Temp_MGQDold"), c.arr("Temp_MGQDold2"), (int)nc, c.p.conv_type,
                    c.p.conv_lo, mg_its, &tn, &tr, &tc);
        convergence(c, c.arr("MGQD_E_avg"), c.arr("MGQD_E_avg_MGQDold"), c.arr("MGQD_E_avg_MGQDold2"), (int)nc,
                    c.p.conv_type, c.p.conv_lo, mg_its, &en, &er, &ec);
        mg_conv = tc && ec;
        mlog.tnorm = tn; mlog.trho = tr; mlog.enorm = en; mlog.erho = er;
        cp(c, "MGQD_E_avg_MGQDold2", "MGQD_E_avg_MGQDold");   // AL:797-800
        cp(c, "MGQD_E_avg_MGQDold", "MGQD_E_avg");
        cp(c, "Temp_MGQDold2", "Temp_MGQDold");
        cp(c, "Temp_MGQDold", "Temp");
    }
    st.mg_total += mg_its;
    rlog.mg = mg_its;
    rlog.kits = c.last_kits;
    int tc = 0, ec = 0;
    double tn, tr, en, er;
    convergence(c, c.arr("Temp"), c.arr("Temp_RTold"), c.arr("Temp_RTold2"), (int)nc, c.p.conv_type, c.p.conv_ho,
                st.rt_its, &tn, &tr, &tc);
    convergence(c, c.arr("HO_E_avg"), c.arr("HO_E_avg_RTold"), c.arr("HO_E_avg_RTold2"), (int)nc, c.p.conv_type,
                c.p.conv_ho, st.rt_its, &en, &er, &ec);
    st.rt_conv = tc && ec;
    rlog.tnorm = tn; rlog.trho = tr; rlog.enorm = en; rlog.erho = er;
    cp(c, "HO_E_avg_RTold2", "HO_E_avg_RTold");   // AL:834-837
    cp(c, "HO_E_avg_RTold", "HO_E_avg");
    cp(c, "Temp_RTold2", "Temp_RTold");
    cp(c, "Temp_RTold", "Temp");
}

int step_end(Ctx &c, const StepState &st, int its_out[4])
{
    TRT_CHECK(hipStreamSynchronize(c.stream));
    if (its_out) { its_out[0] = st.rt_its; its_out[1] = st.mg_total; its_out[2] = st.newt_total; its_out[3] = st.sweeps; }
    if (c.verbose)
        fprintf(stderr, "[step] RT %d, MLOQD %d, Newton %d; direct solves so far: %lld factorizations, %lld on reused "
                "factors (%lld refinement steps)\n", st.rt_its, st.mg_total, st.newt_total, c.nd_factor, c.nd_reused,
                c.nd_reuse_steps);
    return st.rt_its;
}

int step(Ctx &c, int its_out[4], int ho_form)
{
    StepState st;
    step_begin(c, st, ho_form);
    while (!step_done(st)) step_rt_iteration(c, st, false);
    return step_end(c, st, its_out);
}

// The low-order state a full step allocates on first use -- the MLOQD and grey QD
// systems, the MLOQD scratch and the nested-dissection plans and fronts -- allocated
// now (a request loop admits a solve only when all of it fits).
void prepare_solves(Ctx &c)
{
    if (c.mg.nsys == 0) qd_alloc(c, c.mg, c.ng);
    if (c.grey.nsys == 0) qd_alloc(c, c.grey, 1);
    ml_scratch(c);
    if (c.qd_solver == 0 || c.qd_solver == 3) {
        qd_nd_prepare(c, c.mg);
        qd_nd_prepare(c, c.grey);
    }
}

// Full time steps of every member of a batch.  The members step in lockstep rounds of
// RT iterations: each round's sweeps (of the members whose RT iteration starts with one)
// run as one batched persistent launch, then every member still iterating runs the rest
// of its RT iteration -- QD factors, the MLOQD and grey Newton iterations, the
// convergence checks -- on its own stream from a host thread of its own, so the members'
// low-order solves overlap on the device.  A member's arithmetic is that of step():
// its results are the same alone and in the batch.
void batch_step(PkBatch &b, std::vector<int> *its)
{
    const std::vector<Ctx *> &cs = pk_batch_ctxs(b);
    const int n = (int)cs.size();
    for (Ctx *c : cs)
        if (c->pk == nullptr) throw Error("batched steps: a member has no tile engine state");
    std::vector<StepState> st(n);
    for (int i = 0; i < n; ++i) step_begin(*cs[i], st[i], 0);
    std::vector<char> sweep(n);
    std::vector<std::string> errs(n);
    for (;;) {
        bool any = false, anysweep = false;
        for (int i = 0; i < n; ++i) {
            any = any || !step_done(st[i]);
            sweep[i] = step_sweeps_next(st[i]);
            anysweep = anysweep || sweep[i];
        }
        if (!any) break;
        if (anysweep) pk_batch_sweep(b, 3, &sweep);
        std::vector<std::thread> th;
        for (int i = 0; i < n; ++i) {
            if (step_done(st[i])) continue;
            th.emplace_back([&, i]() {
                try {
                    step_rt_iteration(*cs[i], st[i], sweep[i] != 0);
                } catch (const Error &e) {
                    errs[i] = e.msg;
                } catch (const std::exception &e) {
                    errs[i] = e.what();
                }
            });
        }
        for (std::thread &t : th) t.join();
        for (int i = 0; i < n; ++i)
            if (!errs[i].empty()) throw Error("batched step, member " + std::to_string(i) + ": " + errs[i]);
        if (anysweep) pk_batch_check(b);
    }
    if (its) its->assign((size_t)4 * n, 0);
    for (int i = 0; i < n; ++i) step_end(*cs[i], st[i], its ? its->data() + 4 * i : nullptr);
}

// ---------------------------------------------------------------------------
// Tgen_QDf, ALGORITHMS.f08:1031-1095: one sweep per step with Temp prescribed
// (MATERIAL_UPDATE :1042, I_crn_old = I_crn :1046, TRANSPORT_SCB :1047,
// COLLAPSE_INTENSITIES :1056, fg_Calc :1083, Cg_Calc :1095), no low-order solve.
// ---------------------------------------------------------------------------
void tgen_pre(Ctx &c)
{
    c.time += c.p.dt;
    material_update(c);
    c.iold = c.icur;
}

void tgen_post(Ctx &c)
{
    collapse_with_fg(c);
    cg_calc(c);
}

void tgen_step(Ctx &c)
{
    tgen_pre(c);
    transport_sweep(c);
    tgen_post(c);
}

}  // namespace trt
