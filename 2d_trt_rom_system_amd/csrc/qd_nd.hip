// Direct solve of the reduced QD systems by nested dissection (multifrontal LU).
//
// QD_FV's matrix (MainSrc/QD_SOLVERS.f08:40-538) couples a cell's E_avg to its
// four faces, and a face's momentum equation to the two cells beside it and
// the faces of those cells across the other direction.  On this staggered grid
// a line of face unknowns separates the grid: the E_edgV of one face column
// splits a rectangle of cells into left and right parts with no equation
// coupling them except through that line, and likewise a row of E_edgH.  The
// grid is bisected recursively (the longer side first) down to leaves of at
// most 4x4 cells; every tree node owns its separator (a leaf owns all of its
// unknowns), and its front holds those unknowns (I) and the separator unknowns
// on the perimeter of its rectangle (B).
//
// Numeric phase, bottom-up by node height, batched over the nodes of a height
// and the systems solved (every equation scaled by 1 / its largest
// coefficient; iterative refinement with the stored factors at the end while
// the backward error is above 1e-14): assemble the node's matrix entries and right-hand
// side into a dense front, add its children's Schur complements (each child
// in turn, so the sums are deterministic), factor the I block (LU with partial
// pivoting inside I), form A_II^-1 A_IB and the Schur complement
// A_BB - A_BI A_II^-1 A_IB for the parent, and eliminate the right-hand side
// the same way.  Back substitution runs top-down: x_I = A_II^-1 b_I - (A_II^-1
// A_IB) x_B.  Fronts of up to 64 unknowns (the lowest levels) are done by
// nd_front_small, one wavefront each; rocSOLVER getrf/getrs and rocBLAS gemm do
// the dense work on the larger fronts.  Fronts are padded to the largest of
// their height (padding: identity in I, zeros in B).  At 256x256 cells this is
// about 1.1e9 flop per system and fronts of at most 384 unknowns.
#include "qd_view.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <functional>

namespace trt {

#define ND_RB(expr)                                                                        \
    do {                                                                                   \
        rocblas_status _s = (expr);                                                        \
        if (_s != rocblas_status_success)                                                  \
            throw Error(std::string(#expr) + ": rocBLAS status " + std::to_string((int)_s)); \
    } while (0)

namespace {

constexpr int ND_LEAF = 4;      // leaves: at most ND_LEAF x ND_LEAF cells
constexpr int ND_MAXT = 8;      // terms per equation (at most 7)

struct NDLevDev {
    int nn, nIp, nBp, nF;
    double *F, *f;              // fronts [node][sys][nF][nF] (column-major), right-hand sides [node][sys][nF]
    const int *up;              // [node][nBp] position of each B unknown in the parent's front
};

struct NDLevel {
    int nn = 0, nIp = 0, nBp = 0, nF = 0;
    int *rows = nullptr;        // [node][nF] global unknown of each front row (-1: padding)
    int *terms = nullptr;       // [node][nF][ND_MAXT] front column of each term of the row's equation (-1: not here)
    int *up = nullptr;          // [node][nBp]
    int *kids = nullptr;        // [node][2] child as level << 20 | index (-1: none)
    size_t foff = 0, voff = 0, poff = 0;   // offsets of the level in the F, f and pivot buffers (per system chunk)
};

__global__ void nd_zero(double *p, size_t n)
{
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) p[k] = 0.0;
}

// Row equilibration: each equation is scaled by 1 / max |coefficient|, so that
// the partial pivoting inside a front compares rows of one scale (energy
// balances of optically thick cells otherwise dwarf the momentum rows).
__global__ void nd_rowscale(QDView V, const int *sysmap, int ns, double *W)
{
    const size_t tot = (size_t)ns * V.nrow;
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < tot; k += (size_t)gridDim.x * blockDim.x) {
        const int s = (int)(k / V.nrow);
        const size_t row = k - (size_t)s * V.nrow;
        double m = 0.0, dg, rhs;
        qd_row_visit(V, sysmap[s], row, [&](int, int, int, double a) { m = fmax(m, fabs(a)); }, dg, rhs);
        W[k] = m > 0.0 ? 1.0 / m : 1.0;
    }
}

// one block per (node, system): the node's own matrix entries and right-hand side
__global__ void nd_assemble(QDView V, const int *sysmap, int ns, int nIp, int nF, const int *rows, const int *terms,
                            const double *W, double *F, double *f)
{
    const int n = blockIdx.x / ns, s = blockIdx.x % ns;
    double *Fb = F + ((size_t)n * ns + s) * nF * nF;
    double *fb = f + ((size_t)n * ns + s) * nF;
    const int sys = sysmap[s];
    for (int p = threadIdx.x; p < nF; p += blockDim.x) {
        const int r = rows[(size_t)n * nF + p];
        if (r < 0) {
            if (p < nIp) Fb[(size_t)p + (size_t)p * nF] = 1.0;
            continue;
        }
        const int *tc = terms + ((size_t)n * nF + p) * ND_MAXT;
        const double w = W[(size_t)s * V.nrow + r];
        int t = 0;
        double dg, rhs;
        qd_row_visit(V, sys, (size_t)r, [&](int, int, int, double a) {
            const int col = tc[t++];
            if (col >= 0) Fb[(size_t)p + (size_t)col * nF] += w * a;
        }, dg, rhs);
        if (p < nIp) fb[p] += w * rhs;
    }
}

// Child k (0 or 1) of every node of level L: its Schur complement and eliminated
// right-hand side added into the parent's front.  blockIdx.x = (node, system); the
// blocks of blockIdx.y and their threads stride over the child's border block, whose
// entries go to distinct parent entries (up is one-to-one), so no two threads of a
// launch write the same value; the two children are two launches, in a fixed order, so
// each parent entry is summed in the same order as one child after the other.  Each
// thread keeps four updates in flight (the front pointers do not alias).
__global__ __launch_bounds__(256) void nd_pull_child(const NDLevDev *lev, int L, int ns, const int *kids, int k)
{
    const NDLevDev P = lev[L];
    const int n = blockIdx.x / ns, s = blockIdx.x % ns;
    const int kid = kids[n * 2 + k];
    if (kid < 0) return;
    const NDLevDev C = lev[kid >> 20];
    const int ci = kid & 0xfffff;
    double *__restrict__ Fp = P.F + ((size_t)n * ns + s) * P.nF * P.nF;
    const double *__restrict__ Fc = C.F + ((size_t)ci * ns + s) * C.nF * C.nF + (size_t)C.nIp * C.nF + C.nIp;
    const int *__restrict__ up = C.up + (size_t)ci * C.nBp;
    const int nb = C.nBp, tot = nb * nb, nFp = P.nF, nFc = C.nF;
    const int stride = gridDim.y * blockDim.x;
    int e = blockIdx.y * blockDim.x + threadIdx.x;
    for (; e + 3 * stride < tot; e += 4 * stride) {
        int pos[4];
        double val[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int eu = e + u * stride, i = eu % nb, j = eu / nb;
            const int pi = up[i], pj = up[j];
            pos[u] = (pi >= 0 && pj >= 0) ? pi + pj * nFp : -1;
            val[u] = Fc[(size_t)i + (size_t)j * nFc];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (pos[u] >= 0) Fp[pos[u]] += val[u];
    }
    for (; e < tot; e += stride) {
        const int i = e % nb, j = e / nb;
        const int pi = up[i], pj = up[j];
        if (pi >= 0 && pj >= 0) Fp[(size_t)pi + (size_t)pj * nFp] += Fc[(size_t)i + (size_t)j * nFc];
    }
    if (blockIdx.y == 0) {
        double *fp = P.f + ((size_t)n * ns + s) * P.nF;
        const double *fc = C.f + ((size_t)ci * ns + s) * C.nF;
        for (int i = threadIdx.x; i < nb; i += blockDim.x) {
            const int pi = up[i];
            if (pi >= 0) fp[pi] += fc[C.nIp + i];
        }
    }
}

// back substitution helpers: x_B into the front's right-hand side slot, x_I out
__global__ void nd_gather(const double *x, size_t nrow, const int *sysmap, int ns, int nIp, int nBp, int nF,
                          const int *rows, double *f)
{
    const int n = blockIdx.x / ns, s = blockIdx.x % ns;
    double *fb = f + ((size_t)n * ns + s) * nF;
    const double *xs = x + (size_t)sysmap[s] * nrow;
    for (int q = threadIdx.x; q < nBp; q += blockDim.x) {
        const int r = rows[(size_t)n * nF + nIp + q];
        fb[nIp + q] = r >= 0 ? xs[r] : 0.0;
    }
}

__global__ void nd_scatter(double *x, size_t nrow, const int *sysmap, int ns, int nIp, int nF, const int *rows,
                           const double *f)
{
    const int n = blockIdx.x / ns, s = blockIdx.x % ns;
    const double *fb = f + ((size_t)n * ns + s) * nF;
    double *xs = x + (size_t)sysmap[s] * nrow;
    for (int p = threadIdx.x; p < nIp; p += blockDim.x) {
        const int r = rows[(size_t)n * nF + p];
        if (r >= 0) xs[r] = fb[p];
    }
}

// Residual R = b - A x of each system of the chunk (blockIdx.y) together with
// per-block partial sums, in the equilibrated rows (row scales W), of |r|^2 and
// (|A||x| + |b|)^2: the pieces of the normwise backward error.
constexpr int ND_RB_BLOCKS = 64;

__global__ __launch_bounds__(256) void nd_residual_bkwd(QDView V, const int *sysmap, const double *x, const double *W,
                                                        double *R, double *part)
{
    __shared__ double sh[2][4];
    const int s = blockIdx.y;
    const int sys = sysmap[s];
    const double *xs = x + (size_t)sys * V.nrow;
    const size_t off = (size_t)s * V.nrow;
    double rr = 0.0, ss = 0.0;
    for (size_t row = blockIdx.x * (size_t)blockDim.x + threadIdx.x; row < V.nrow;
         row += (size_t)gridDim.x * blockDim.x) {
        double acc = 0.0, aax = 0.0, dg, rhs;
        qd_row_visit(V, sys, row, [&](int kind, int i, int j, double a) {
            const double t = a * xs[qd_index(V, kind, i, j)];
            acc += t;
            aax += fabs(t);
        }, dg, rhs);
        const double r = rhs - acc, w = W[off + row];
        R[off + row] = r;
        rr += (w * r) * (w * r);
        ss += (w * (aax + fabs(rhs))) * (w * (aax + fabs(rhs)));
    }
    for (int o = 32; o > 0; o >>= 1) {
        rr += __shfl_xor(rr, o, 64);
        ss += __shfl_xor(ss, o, 64);
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sh[0][wv] = rr;
        sh[1][wv] = ss;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
            a += sh[0][k];
            b += sh[1][k];
        }
        part[((size_t)s * gridDim.x + blockIdx.x) * 2] = a;
        part[((size_t)s * gridDim.x + blockIdx.x) * 2 + 1] = b;
    }
}

// backward error of each system of the chunk from the partial sums
__global__ void nd_bkwd_fin(const double *part, int nblk, int ns, double *out)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    double a = 0.0, b = 0.0;
    for (int k = 0; k < nblk; ++k) {
        a += part[((size_t)s * nblk + k) * 2];
        b += part[((size_t)s * nblk + k) * 2 + 1];
    }
    out[s] = a == 0.0 ? 0.0 : sqrt(a) / sqrt(b);
}

// iterative refinement: right-hand sides of the fronts from the residual, the
// children's eliminated parts, and x += dx
__global__ void nd_rhs(const double *R, const double *W, size_t nrow, int ns, int nIp, int nF, const int *rows,
                       double *f)
{
    const int n = blockIdx.x / ns, s = blockIdx.x % ns;
    double *fb = f + ((size_t)n * ns + s) * nF;
    for (int p = threadIdx.x; p < nIp; p += blockDim.x) {
        const int r = rows[(size_t)n * nF + p];
        fb[p] = r >= 0 ? W[(size_t)s * nrow + r] * R[(size_t)s * nrow + r] : 0.0;
    }
}

__global__ void nd_pull_vec(const NDLevDev *lev, int L, int ns, const int *kids)
{
    const NDLevDev P = lev[L];
    const int n = blockIdx.x / ns, s = blockIdx.x % ns;
    double *fp = P.f + ((size_t)n * ns + s) * P.nF;
    for (int k = 0; k < 2; ++k) {
        const int kid = kids[n * 2 + k];
        if (kid < 0) continue;
        const NDLevDev C = lev[kid >> 20];
        const int ci = kid & 0xfffff;
        const double *fc = C.f + ((size_t)ci * ns + s) * C.nF;
        const int *up = C.up + (size_t)ci * C.nBp;
        for (int i = threadIdx.x; i < C.nBp; i += blockDim.x) {
            const int pi = up[i];
            if (pi >= 0) fp[pi] += fc[C.nIp + i];
        }
        __syncthreads();
    }
}

// Right-hand side of a front in a refinement solve, as nd_rhs and nd_pull_vec leave
// it, inside the level's forward kernel: the pivot part of W R, then the children's
// eliminated right-hand sides, one child after the other (the same additions in the
// same order).  The block is the front of (node, system) blockIdx.x.
struct NDFwdIn {
    const double *R, *W;
    size_t nrow;
    const int *rows, *kids;
    const NDLevDev *lev;
    int L, ns;
};

__device__ __forceinline__ void nd_fwd_in(const NDFwdIn &a, int nIp, int nF, double *fb)
{
    const int n = blockIdx.x / a.ns, s = blockIdx.x % a.ns;
    for (int p = threadIdx.x; p < nIp; p += blockDim.x) {
        const int r = a.rows[(size_t)n * nF + p];
        fb[p] = r >= 0 ? a.W[(size_t)s * a.nrow + r] * a.R[(size_t)s * a.nrow + r] : 0.0;
    }
    __syncthreads();
    if (a.L == 0) return;
    for (int k = 0; k < 2; ++k) {
        const int kid = a.kids[n * 2 + k];
        if (kid >= 0) {
            const NDLevDev C = a.lev[kid >> 20];
            const int ci = kid & 0xfffff;
            const double *fc = C.f + ((size_t)ci * a.ns + s) * C.nF;
            const int *up = C.up + (size_t)ci * C.nBp;
            for (int i = threadIdx.x; i < C.nBp; i += blockDim.x) {
                const int pi = up[i];
                if (pi >= 0) fb[pi] += fc[C.nIp + i];
            }
        }
        __syncthreads();
    }
}

// The solve phase on fronts whose pivot block holds its explicit inverse G = A_II^-1
// (the factor-reuse mode, levels of fronts too large for nd_front_small), one block
// per (node, system) of a level.  Forward (after nd_fwd_in):
// z_I = G f_I in place of f_I, then f_B -= F_BI z_I.  Backward, with x_B read from the
// ancestors' solution: x_I = z_I - (A_II^-1 F_IB) x_B, written to the solution.  Matrix
// columns are read across the threads (column-major fronts): every load instruction is
// coalesced.
__global__ __launch_bounds__(256) void nd_fwd_inv(const double *F, double *f, int nIp, int nBp, int nF, NDFwdIn in)
{
    extern __shared__ double sh[];            // f_I, then z
    const double *Fb = F + (size_t)blockIdx.x * nF * nF;
    double *fb = f + (size_t)blockIdx.x * nF;
    nd_fwd_in(in, nIp, nF, fb);
    for (int i = threadIdx.x; i < nIp; i += blockDim.x) sh[i] = fb[i];
    __syncthreads();
    for (int i = threadIdx.x; i < nIp; i += blockDim.x) {
        double a0 = 0.0, a1 = 0.0;                // two chains: more loads in flight
        int j = 0;
#pragma unroll 4
        for (; j + 1 < nIp; j += 2) {
            a0 = fma(Fb[i + (size_t)j * nF], sh[j], a0);
            a1 = fma(Fb[i + (size_t)(j + 1) * nF], sh[j + 1], a1);
        }
        if (j < nIp) a0 = fma(Fb[i + (size_t)j * nF], sh[j], a0);
        const double acc = a0 + a1;
        sh[nIp + i] = acc;
        fb[i] = acc;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < nBp; q += blockDim.x) {
        double a0 = fb[nIp + q], a1 = 0.0;
        int j = 0;
#pragma unroll 4
        for (; j + 1 < nIp; j += 2) {
            a0 = fma(-Fb[nIp + q + (size_t)j * nF], sh[nIp + j], a0);
            a1 = fma(-Fb[nIp + q + (size_t)(j + 1) * nF], sh[nIp + j + 1], a1);
        }
        if (j < nIp) a0 = fma(-Fb[nIp + q + (size_t)j * nF], sh[nIp + j], a0);
        fb[nIp + q] = a0 + a1;
    }
}

// backward, the fused form of nd_gather, the back substitution and nd_scatter: x_B from
// the solution D ([system][nrow], written by the ancestors' launches), x_I into D
__global__ __launch_bounds__(256) void nd_bwd_inv(const double *F, const double *f, int nIp, int nBp, int nF,
                                                  double *D, size_t nrow, int ns, const int *rows)
{
    extern __shared__ double sh[];            // x_B
    const int n = blockIdx.x / ns, s = blockIdx.x % ns;
    const double *Fb = F + (size_t)blockIdx.x * nF * nF;
    const double *fb = f + (size_t)blockIdx.x * nF;
    const int *rw = rows + (size_t)n * nF;
    double *Ds = D + (size_t)s * nrow;
    for (int q = threadIdx.x; q < nBp; q += blockDim.x) {
        const int r = rw[nIp + q];
        sh[q] = r >= 0 ? Ds[r] : 0.0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nIp; i += blockDim.x) {
        double a0 = fb[i], a1 = 0.0;
        int q = 0;
#pragma unroll 4
        for (; q + 1 < nBp; q += 2) {
            a0 = fma(-Fb[i + (size_t)(nIp + q) * nF], sh[q], a0);
            a1 = fma(-Fb[i + (size_t)(nIp + q + 1) * nF], sh[q + 1], a1);
        }
        if (q < nBp) a0 = fma(-Fb[i + (size_t)(nIp + q) * nF], sh[q], a0);
        const int r = rw[i];
        if (r >= 0) Ds[r] = a0 + a1;
    }
}

__global__ void nd_add(double *x, const double *D, const int *sysmap, int ns, size_t nrow)
{
    const size_t tot = (size_t)ns * nrow;
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < tot; k += (size_t)gridDim.x * blockDim.x) {
        const int s = (int)(k / nrow);
        x[(size_t)sysmap[s] * nrow + (k - (size_t)s * nrow)] += D[k];
    }
}

// Fronts of at most ND_SF unknowns with at most ND_SI pivot rows, one wavefront
// each (lane t holds row t of the front in registers): partial pivoting among
// the pivot rows, elimination of the pivot columns from every row (pivot and
// border rows, and the right-hand side f), then back substitution of the border
// columns and f.  The front is left as getrf + getrs + gemm leave it: LU and the
// getrf pivots in the pivot block (rows in pivot order), A_II^-1 F_IB and
// A_II^-1 f_I, F_BI unchanged, the Schur complement F_BB - F_BI A_II^-1 F_IB and
// f_B - F_BI A_II^-1 f_I; info = first zero pivot (1-based) as getrf reports.
constexpr int ND_SF = 64, ND_SI = 48;

struct NDSmall {
    double r[ND_SF];       // this lane's row of the front
    double rf;             // its right-hand-side entry
    double *prow;          // LDS: the current pivot row, f entry last
    int *s_orig;           // LDS: front row chosen at each pivot step
    int t, nF, nIp, pos, zero_at;
    bool live, pivot_row, chosen;
};

// NIP > 0 fixes the front's pivot count at compile time: the columns past nF are
// loaded as zeros and eliminated like the others (they stay zero and are never
// stored), so no column needs a runtime guard and a kernel is about a tenth of the
// code of the NIP = 0 form, which guards every column by the runtime nIp and nF
// (that form's 426 KB of code streams through the 64 KB instruction cache once per
// wave).  Both forms do the same operations on every stored entry, bitwise.
template <int NIP>
__device__ __forceinline__ bool nd_col(const NDSmall &w, int j)
{
    return NIP > 0 || j < w.nF;
}
template <int NIP>
__device__ __forceinline__ int nd_nip(const NDSmall &w)
{
    return NIP > 0 ? NIP : w.nIp;
}

// One step of the pivot search's all-reduce: the partner lane's candidate by a DPP move
// within a row of 16 lanes (CTRL: 0xB1 and 0x4E swap neighbours and pairs in a quad,
// 0x141 mirrors 8 lanes, 0x140 mirrors 16), or by a lane shuffle across rows (XOR 16, 32).
// Every step pairs each lane with one of another group, so after the six steps every lane
// holds the largest candidate; the order (largest |entry|, then lowest lane) is total, so
// the choice does not depend on the network.
template <int CTRL>
__device__ __forceinline__ void nd_pivot_step(double &key, int &who)
{
    double ok;
    int ow;
    if constexpr (CTRL < 0x10) {
        // CTRL 1 and 2: lanes t ^ 16 and t ^ 32
        ok = __shfl_xor(key, CTRL << 4);
        ow = __shfl_xor(who, CTRL << 4);
    } else {
        const long long kb = __double_as_longlong(key);
        const int lo = __builtin_amdgcn_update_dpp(0, (int)kb, CTRL, 0xF, 0xF, false);
        const int hi = __builtin_amdgcn_update_dpp(0, (int)(kb >> 32), CTRL, 0xF, 0xF, false);
        ok = __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
        ow = __builtin_amdgcn_update_dpp(0, who, CTRL, 0xF, 0xF, false);
    }
    if (ok > key || (ok == key && ow < who)) {
        key = ok;
        who = ow;
    }
}

// pivot step K (the template argument keeps every register index constant)
template <int K, int NIP>
__device__ __forceinline__ void nd_small_elim(NDSmall &w)
{
    if constexpr (K < (NIP > 0 ? NIP : ND_SI)) {
        if (NIP > 0 || K < w.nIp) {
            // the largest |entry| of column K among the pivot rows not chosen yet
            double key = (w.pivot_row && !w.chosen) ? fabs(w.r[K]) : -1.0;
            int who = w.t;
            if constexpr (NIP > 0) {
                nd_pivot_step<0xB1>(key, who);
                nd_pivot_step<0x4E>(key, who);
                nd_pivot_step<0x141>(key, who);
                nd_pivot_step<0x140>(key, who);
                nd_pivot_step<1>(key, who);
                nd_pivot_step<2>(key, who);
            } else {
                for (int o = 1; o < 64; o <<= 1) {
                    const double ok = __shfl_xor(key, o);
                    const int ow = __shfl_xor(who, o);
                    if (ok > key || (ok == key && ow < who)) {
                        key = ok;
                        who = ow;
                    }
                }
            }
            const int p = __builtin_amdgcn_readfirstlane(who);
            if (w.t == p) {
                w.chosen = true;
                w.pos = K;
                w.s_orig[K] = w.t;
#pragma unroll
                for (int j = K; j < ND_SF; ++j)
                    if (nd_col<NIP>(w, j)) w.prow[j] = w.r[j];
                w.prow[ND_SF] = w.rf;
            }
            __syncthreads();
            const double piv = w.prow[K];
            if (piv == 0.0 && w.zero_at == 0) w.zero_at = K + 1;
            if (w.live && !w.chosen) {
                const double l = piv != 0.0 ? w.r[K] * (1.0 / piv) : 0.0;
                w.r[K] = l;
#pragma unroll
                for (int j = K + 1; j < ND_SF; ++j)
                    if (nd_col<NIP>(w, j)) w.r[j] = fma(-l, w.prow[j], w.r[j]);
                w.rf = fma(-l, w.prow[ND_SF], w.rf);
            }
            __syncthreads();
            nd_small_elim<K + 1, NIP>(w);
        }
    }
}

// back substitution step K (last pivot first): x_K = z_K / U_KK, then the
// pivot rows chosen before K remove U_{row,K} x_K
template <int K, int NIP>
__device__ __forceinline__ void nd_small_back(NDSmall &w)
{
    if constexpr (K >= 0) {
        if (NIP > 0 || K < w.nIp) {
            if (w.pos == K) {
                const double inv = w.r[K] != 0.0 ? 1.0 / w.r[K] : 0.0;
#pragma unroll
                for (int j = NIP; j < ND_SF; ++j)
                    if (j >= nd_nip<NIP>(w) && nd_col<NIP>(w, j)) {
                        w.r[j] *= inv;
                        w.prow[j] = w.r[j];
                    }
                w.rf *= inv;
                w.prow[ND_SF] = w.rf;
            }
            __syncthreads();
            if (w.pivot_row && w.pos < K) {
                const double u = w.r[K];
#pragma unroll
                for (int j = NIP; j < ND_SF; ++j)
                    if (j >= nd_nip<NIP>(w) && nd_col<NIP>(w, j)) w.r[j] = fma(-u, w.prow[j], w.r[j]);
                w.rf = fma(-u, w.prow[ND_SF], w.rf);
            }
            __syncthreads();
        }
        nd_small_back<K - 1, NIP>(w);
    }
}

template <int NIP>
__global__ __launch_bounds__(64) void nd_front_small(double *F, double *f, int nF, int nIp, size_t sF, size_t sf,
                                                     rocblas_int *ipiv, rocblas_int *info, int *perm, int nb)
{
    __shared__ double prow[ND_SF + 1];
    __shared__ int s_orig[ND_SI];
    __shared__ int s_arr[ND_SI], s_where[ND_SI];
    const size_t fr = blockIdx.x;
    if (fr >= (size_t)nb) return;
    double *A = F + fr * sF;
    double *v = f + fr * sf;
    NDSmall w;
    w.t = threadIdx.x;
    w.nF = nF;
    w.nIp = nIp;
    w.prow = prow;
    w.s_orig = s_orig;
    w.live = w.t < nF;
    w.pivot_row = w.t < nIp;
    w.chosen = false;
    w.pos = -1;
    w.zero_at = 0;
#pragma unroll
    for (int j = 0; j < ND_SF; ++j) w.r[j] = (w.live && j < nF) ? A[(size_t)j * nF + w.t] : 0.0;
    w.rf = w.live ? v[w.t] : 0.0;
    nd_small_elim<0, NIP>(w);
    nd_small_back<(NIP > 0 ? NIP : ND_SI) - 1, NIP>(w);
    if (w.pivot_row) {
        // LU row and A_II^-1 F_IB row at the row's pivot position
#pragma unroll
        for (int j = 0; j < ND_SF; ++j)
            if (j < nF) A[(size_t)j * nF + w.pos] = w.r[j];
        v[w.pos] = w.rf;
    } else if (w.live) {
        // border row: F_BI stays, Schur complement and reduced f_B
#pragma unroll
        for (int j = 0; j < ND_SF; ++j)
            if (j >= nIp && j < nF) A[(size_t)j * nF + w.t] = w.r[j];
        v[w.t] = w.rf;
    }
    // the front row chosen at each pivot step (the row permutation, for nd_fwd_lu_small)
    if (w.t < nIp) perm[fr * nIp + w.t] = s_orig[w.t];
    // getrf's pivot sequence: at step k rows k and ipiv[k]-1 of the current order swap
    if (w.t == 0) {
        rocblas_int *pv = ipiv + fr * nIp;
        for (int i = 0; i < nIp; ++i) {
            s_arr[i] = i;
            s_where[i] = i;
        }
        for (int k = 0; k < nIp; ++k) {
            const int o = s_orig[k], pp = s_where[o], ak = s_arr[k];
            pv[k] = pp + 1;
            s_arr[pp] = ak;
            s_where[ak] = pp;
            s_arr[k] = o;
            s_where[o] = k;
        }
        info[fr] = w.zero_at;
    }
}

// The forward solve phase on a level of small fronts with the LU factors nd_front_small
// left (one wavefront per front, lane t holding row t of the front in registers, all of
// it loaded at once): z_I = A_II^-1 f_I by the row permutation, the unit lower and the
// upper triangular solves, then f_B -= F_BI z_I, the pivots' values broadcast by
// v_readlane (the source lane is uniform, so no LDS crossbar round trip).  The template
// steps keep every register index constant.
__device__ __forceinline__ double nd_lane(double x, int lane)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)b, lane), hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

template <int K>
__device__ __forceinline__ void nd_lu_fwd_step(const double (&r)[ND_SI], double &y, int t, int nIp)
{
    if constexpr (K < ND_SI) {
        if (K < nIp) {
            const double yk = nd_lane(y, K);
            if (t > K && t < nIp) y = fma(-r[K], yk, y);
            nd_lu_fwd_step<K + 1>(r, y, t, nIp);
        }
    }
}

template <int K>
__device__ __forceinline__ void nd_lu_bwd_step(const double (&r)[ND_SI], double &y, int t, int nIp)
{
    if constexpr (K >= 0) {
        if (K < nIp) {
            if (t == K) y = y / r[K];
            const double yk = nd_lane(y, K);
            if (t < K) y = fma(-r[K], yk, y);
        }
        nd_lu_bwd_step<K - 1>(r, y, t, nIp);
    }
}

__global__ __launch_bounds__(64) void nd_fwd_lu_small(const double *F, double *f, const int *perm, int nIp, int nBp,
                                                      int nF, NDFwdIn in)
{
    const size_t fr = blockIdx.x;
    const double *A = F + fr * nF * nF;
    double *v = f + fr * nF;
    const int t = threadIdx.x;
    nd_fwd_in(in, nIp, nF, v);
    const bool live = t < nF;
    double r[ND_SI];                              // row t: L and U (pivot rows) or F_BI (border rows)
#pragma unroll
    for (int k = 0; k < ND_SI; ++k) r[k] = (live && k < nIp) ? A[t + (size_t)k * nF] : 0.0;
    double y = t < nIp ? v[perm[fr * nIp + t]] : (live ? v[t] : 0.0);
    double z = t < nIp ? y : 0.0;
    nd_lu_fwd_step<0>(r, z, t, nIp);
    nd_lu_bwd_step<ND_SI - 1>(r, z, t, nIp);
    // border rows: f_B -= F_BI z_I
#pragma unroll
    for (int k = 0; k < ND_SI; ++k) {
        if (k < nIp) {
            const double zk = nd_lane(z, k);
            if (t >= nIp) y = fma(-r[k], zk, y);
        }
    }
    if (t < nIp) v[t] = z;
    else if (live) v[t] = y;
}

__global__ void nd_info_max(const rocblas_int *info, size_t n, int *out)
{
    int m = 0;
    for (size_t k = threadIdx.x; k < n; k += blockDim.x) m = max(m, (int)info[k]);
    for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
    __shared__ int sh[16];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t = max(t, sh[k]);
        *out = t;
    }
}

int nblocks(size_t n) { return (int)std::min<size_t>(std::max<size_t>((n + 255) / 256, 1), 65536); }

}  // namespace

struct NDPlan {
    int nx = 0, ny = 0;
    size_t nrow = 0;
    std::vector<NDLevel> lev;
    NDLevDev *levdev = nullptr;     // device copy of the level table for the current chunk
    int cap = 0;                    // systems per chunk
    size_t fsz = 0, vsz = 0, psz = 0;   // per-system doubles of F, f and pivots over all levels
    double *F = nullptr, *f = nullptr;
    rocblas_int *ipiv = nullptr, *info = nullptr;
    int *perm = nullptr;            // small fronts: the front row of each pivot position
    int *sysmap = nullptr, *imax = nullptr;
    int *ident = nullptr;           // 0, 1, ..., cap-1
    double *R = nullptr, *D = nullptr;   // refinement residual and correction [cap][nrow]
    double *W = nullptr;                 // row scales [cap][nrow]
    double *part = nullptr, *bkwd = nullptr;   // backward-error partial sums and results of a chunk
    size_t ninfo = 0;
    rocblas_handle h = nullptr;
    double flops = 0.0;             // per system
    // the factors the last numeric phase left in F, ipiv and W, of the systems
    // fact_which (one chunk): a preconditioner for later solves of those systems
    bool fact_ok = false;
    std::vector<int> fact_which;
    bool ginv = false;              // the pivot blocks of the large fronts hold A_II^-1 instead of their LU factors
};

// Symbolic phase: the dissection tree, fronts, assembly and extend-add maps.
static NDPlan *nd_plan(Ctx &c, const QDSystems &s)
{
    const int nx = s.nx, ny = s.ny;
    const size_t nc = (size_t)nx * ny, nv = (size_t)(nx + 1) * ny;
    auto C = [&](int i, int j) { return (int)(i + (size_t)nx * j); };
    auto Vf = [&](int f, int j) { return (int)(nc + f + (size_t)(nx + 1) * j); };
    auto Hf = [&](int i, int f) { return (int)(nc + nv + i + (size_t)nx * f); };
    struct Node {
        int kid[2] = {-1, -1}, height = 0, lo = 0, level = 0, index = 0, parent = -1;
        std::vector<int> I, B;
    };
    std::vector<Node> nodes;      // post-order
    std::function<int(int, int, int, int)> rec = [&](int i0, int i1, int j0, int j1) -> int {
        Node nd;
        if (i1 - i0 <= ND_LEAF && j1 - j0 <= ND_LEAF) {
            for (int j = j0; j < j1; ++j)
                for (int i = i0; i < i1; ++i) nd.I.push_back(C(i, j));
            for (int j = j0; j < j1; ++j)
                for (int f = i0 + 1; f < i1; ++f) nd.I.push_back(Vf(f, j));
            if (i0 == 0) for (int j = j0; j < j1; ++j) nd.I.push_back(Vf(0, j));
            if (i1 == nx) for (int j = j0; j < j1; ++j) nd.I.push_back(Vf(nx, j));
            for (int f = j0 + 1; f < j1; ++f)
                for (int i = i0; i < i1; ++i) nd.I.push_back(Hf(i, f));
            if (j0 == 0) for (int i = i0; i < i1; ++i) nd.I.push_back(Hf(i, 0));
            if (j1 == ny) for (int i = i0; i < i1; ++i) nd.I.push_back(Hf(i, ny));
        } else if (i1 - i0 >= j1 - j0) {
            const int f = (i0 + i1) / 2;
            nd.kid[0] = rec(i0, f, j0, j1);
            nd.kid[1] = rec(f, i1, j0, j1);
            for (int j = j0; j < j1; ++j) nd.I.push_back(Vf(f, j));
        } else {
            const int f = (j0 + j1) / 2;
            nd.kid[0] = rec(i0, i1, j0, f);
            nd.kid[1] = rec(i0, i1, f, j1);
            for (int i = i0; i < i1; ++i) nd.I.push_back(Hf(i, f));
        }
        if (i0 > 0) for (int j = j0; j < j1; ++j) nd.B.push_back(Vf(i0, j));
        if (i1 < nx) for (int j = j0; j < j1; ++j) nd.B.push_back(Vf(i1, j));
        if (j0 > 0) for (int i = i0; i < i1; ++i) nd.B.push_back(Hf(i, j0));
        if (j1 < ny) for (int i = i0; i < i1; ++i) nd.B.push_back(Hf(i, j1));
        for (int k = 0; k < 2; ++k)
            if (nd.kid[k] >= 0) nd.height = std::max(nd.height, nodes[nd.kid[k]].height + 1);
        nodes.push_back(std::move(nd));
        return (int)nodes.size() - 1;
    };
    const int root = rec(0, nx, 0, ny);
    // subtree ranges [lo, self] in post-order, parents
    std::function<int(int)> span = [&](int k) -> int {
        int lo = k;
        for (int q = 0; q < 2; ++q)
            if (nodes[k].kid[q] >= 0) {
                nodes[nodes[k].kid[q]].parent = k;
                lo = std::min(lo, span(nodes[k].kid[q]));
            }
        nodes[k].lo = lo;
        return lo;
    };
    span(root);
    auto anc_or_self = [&](int a, int k) { return nodes[a].lo <= k && k <= a; };
    std::vector<int> owner(s.nrow, -1);
    for (int k = 0; k < (int)nodes.size(); ++k)
        for (int u : nodes[k].I) owner[u] = k;
    for (size_t u = 0; u < s.nrow; ++u)
        if (owner[u] < 0) throw Error("nested dissection: an unknown has no owner");

    NDPlan *P = new NDPlan();
    P->nx = nx; P->ny = ny; P->nrow = s.nrow;
    const int H = nodes[root].height;
    P->lev.resize(H + 1);
    for (int k = 0; k < (int)nodes.size(); ++k) {
        NDLevel &L = P->lev[nodes[k].height];
        nodes[k].level = nodes[k].height;
        nodes[k].index = L.nn++;
        L.nIp = std::max(L.nIp, (int)nodes[k].I.size());
        L.nBp = std::max(L.nBp, (int)nodes[k].B.size());
    }
    for (NDLevel &L : P->lev) L.nF = L.nIp + L.nBp;
    // front position of a global unknown in node k's front
    auto front_pos = [&](int k, int u) -> int {
        const Node &nd = nodes[k];
        const NDLevel &L = P->lev[nd.level];
        for (size_t p = 0; p < nd.I.size(); ++p) if (nd.I[p] == u) return (int)p;
        for (size_t q = 0; q < nd.B.size(); ++q) if (nd.B[q] == u) return L.nIp + (int)q;
        return -1;
    };
    // host stencil walk: coefficient values are irrelevant, only the terms' unknowns
    QDView V{};
    V.nx = nx; V.ny = ny; V.nb = std::max(nx, ny);
    V.nc = nc; V.nv = nv; V.nh = (size_t)nx * (ny + 1); V.nrow = s.nrow;
    std::vector<double> zc((size_t)NCOEF * nc, 0.0), zb((size_t)NBND * V.nb, 0.0);
    V.coef = zc.data(); V.bnd = zb.data(); V.cs = nullptr;
    std::vector<std::vector<int>> hrows(H + 1), hterms(H + 1), hup(H + 1), hkids(H + 1);
    for (int l = 0; l <= H; ++l) {
        const NDLevel &L = P->lev[l];
        hrows[l].assign((size_t)L.nn * L.nF, -1);
        hterms[l].assign((size_t)L.nn * L.nF * ND_MAXT, -1);
        hup[l].assign((size_t)L.nn * std::max(L.nBp, 1), -1);
        hkids[l].assign((size_t)L.nn * 2, -1);
    }
    double flops = 0.0;
    for (int k = 0; k < (int)nodes.size(); ++k) {
        const Node &nd = nodes[k];
        const int l = nd.level, n = nd.index;
        const NDLevel &L = P->lev[l];
        // a local lookup for the node's front (small fronts: linear search is fine)
        std::vector<std::pair<int, int>> loc;
        for (size_t p = 0; p < nd.I.size(); ++p) loc.push_back({nd.I[p], (int)p});
        for (size_t q = 0; q < nd.B.size(); ++q) loc.push_back({nd.B[q], L.nIp + (int)q});
        std::sort(loc.begin(), loc.end());
        auto pos = [&](int u) {
            auto it = std::lower_bound(loc.begin(), loc.end(), std::make_pair(u, -1));
            return (it != loc.end() && it->first == u) ? it->second : -1;
        };
        auto walk = [&](int p, int r, bool own_row) {
            int t = 0;
            double dg, rhs;
            qd_row_visit(V, 0, (size_t)r, [&](int kind, int i, int j, double) {
                const int col = (int)qd_index(V, kind, i, j);
                const int o = owner[col];
                int fc = -1;
                if (own_row) {
                    if (anc_or_self(o, k)) {
                        fc = pos(col);
                        if (fc < 0) throw Error("nested dissection: a coupling leaves its front");
                    } else if (!anc_or_self(k, o)) {
                        throw Error("nested dissection: a coupling crosses a separator");
                    }
                } else if (o == k) {
                    fc = pos(col);
                }
                if (t >= ND_MAXT) throw Error("nested dissection: too many terms in an equation");
                hterms[l][((size_t)n * L.nF + p) * ND_MAXT + t] = fc;
                ++t;
            }, dg, rhs);
        };
        for (size_t p = 0; p < nd.I.size(); ++p) {
            hrows[l][(size_t)n * L.nF + p] = nd.I[p];
            walk((int)p, nd.I[p], true);
        }
        for (size_t q = 0; q < nd.B.size(); ++q) {
            const int p = L.nIp + (int)q;
            hrows[l][(size_t)n * L.nF + p] = nd.B[q];
            walk(p, nd.B[q], false);
            if (nd.parent >= 0) {
                const int pp = front_pos(nd.parent, nd.B[q]);
                if (pp < 0) throw Error("nested dissection: a boundary unknown is missing from the parent front");
                hup[l][(size_t)n * L.nBp + q] = pp;
            }
        }
        for (int q = 0; q < 2; ++q)
            if (nd.kid[q] >= 0) hkids[l][(size_t)n * 2 + q] = (nodes[nd.kid[q]].level << 20) | nodes[nd.kid[q]].index;
        const double nI = (double)nd.I.size(), nB = (double)nd.B.size();
        flops += 2.0 / 3.0 * nI * nI * nI + 2.0 * nI * nI * nB + 2.0 * nB * nB * nI;
    }
    P->flops = flops;
    size_t fo = 0, vo = 0, po = 0;
    for (int l = 0; l <= H; ++l) {
        NDLevel &L = P->lev[l];
        L.rows = static_cast<int *>(c.dalloc(hrows[l].size() * sizeof(int)));
        L.terms = static_cast<int *>(c.dalloc(hterms[l].size() * sizeof(int)));
        L.up = static_cast<int *>(c.dalloc(hup[l].size() * sizeof(int)));
        L.kids = static_cast<int *>(c.dalloc(hkids[l].size() * sizeof(int)));
        TRT_CHECK(hipMemcpyAsync(L.rows, hrows[l].data(), hrows[l].size() * sizeof(int), hipMemcpyHostToDevice,
                                 c.stream));
        TRT_CHECK(hipMemcpyAsync(L.terms, hterms[l].data(), hterms[l].size() * sizeof(int), hipMemcpyHostToDevice,
                                 c.stream));
        TRT_CHECK(hipMemcpyAsync(L.up, hup[l].data(), hup[l].size() * sizeof(int), hipMemcpyHostToDevice,
                                 c.stream));
        TRT_CHECK(hipMemcpyAsync(L.kids, hkids[l].data(), hkids[l].size() * sizeof(int), hipMemcpyHostToDevice,
                                 c.stream));
        L.foff = fo; L.voff = vo; L.poff = po;
        fo += (size_t)L.nn * L.nF * L.nF;
        vo += (size_t)L.nn * L.nF;
        po += (size_t)L.nn * L.nIp;
    }
    P->fsz = fo; P->vsz = vo; P->psz = po;
    if (c.verbose > 1)
        for (size_t l = 0; l < P->lev.size(); ++l)
            fprintf(stderr, "[qd_nd plan] level %zu: %d fronts, %d pivots + %d border = %d%s\n", l, P->lev[l].nn,
                    P->lev[l].nIp, P->lev[l].nBp, P->lev[l].nF,
                    P->lev[l].nF <= ND_SF && P->lev[l].nIp <= ND_SI ? " (one wavefront each)" : "");
    // systems per chunk: fronts of all levels within 8 GiB
    const double budget = 8.0 * (1u << 30);
    P->cap = (int)std::max(1.0, std::min<double>(s.nsys, std::floor(budget / (8.0 * (double)fo))));
    P->F = static_cast<double *>(c.dalloc((size_t)P->cap * fo * sizeof(double)));
    P->f = static_cast<double *>(c.dalloc((size_t)P->cap * vo * sizeof(double)));
    P->ipiv = static_cast<rocblas_int *>(c.dalloc((size_t)P->cap * po * sizeof(rocblas_int)));
    P->perm = static_cast<int *>(c.dalloc((size_t)P->cap * po * sizeof(int)));
    size_t ni = 0;
    for (const NDLevel &L : P->lev) ni += (size_t)L.nn;
    P->ninfo = ni * P->cap;
    P->info = static_cast<rocblas_int *>(c.dalloc(P->ninfo * sizeof(rocblas_int)));
    P->sysmap = static_cast<int *>(c.dalloc((size_t)std::max(P->cap, 1) * sizeof(int)));
    P->ident = static_cast<int *>(c.dalloc((size_t)std::max(P->cap, 1) * sizeof(int)));
    {
        std::vector<int> id(std::max(P->cap, 1));
        for (size_t k = 0; k < id.size(); ++k) id[k] = (int)k;
        TRT_CHECK(hipMemcpyAsync(P->ident, id.data(), id.size() * sizeof(int), hipMemcpyHostToDevice, c.stream));
        TRT_CHECK(hipStreamSynchronize(c.stream));   // the plan's tables are uploaded on the stream that uses them
    }
    P->R = static_cast<double *>(c.dalloc((size_t)P->cap * s.nrow * sizeof(double)));
    P->D = static_cast<double *>(c.dalloc((size_t)P->cap * s.nrow * sizeof(double)));
    P->W = static_cast<double *>(c.dalloc((size_t)P->cap * s.nrow * sizeof(double)));
    P->part = static_cast<double *>(c.dalloc((size_t)std::max(P->cap, 1) * ND_RB_BLOCKS * 2 * sizeof(double)));
    P->bkwd = static_cast<double *>(c.dalloc((size_t)std::max(P->cap, 1) * sizeof(double)));
    P->imax = static_cast<int *>(c.dalloc(sizeof(int)));
    P->levdev = static_cast<NDLevDev *>(c.dalloc((H + 1) * sizeof(NDLevDev)));
    ND_RB(rocblas_create_handle(&P->h));
    return P;
}

// fronts that nd_front_small factors (one wavefront each)
static bool nd_small(const NDLevel &L) { return L.nF <= ND_SF && L.nIp <= ND_SI; }

// The small-front kernel for a level: the form compiled for its pivot count where
// there is one (the leaves of 4x4 cells and the separator levels of 4 and 8 pivots
// above them), else the form with runtime sizes; TRT_ND_SMALL_GENERIC=1 takes the latter
// for every level (the two give bitwise equal fronts).
static const void *nd_front_small_kernel(int nIp)
{
    if (const char *v = getenv("TRT_ND_SMALL_GENERIC"); v && atoi(v) != 0)
        return reinterpret_cast<const void *>(&nd_front_small<0>);
    switch (nIp) {
    case 4: return reinterpret_cast<const void *>(&nd_front_small<4>);
    case 8: return reinterpret_cast<const void *>(&nd_front_small<8>);
    case 48: return reinterpret_cast<const void *>(&nd_front_small<48>);
    default: return reinterpret_cast<const void *>(&nd_front_small<0>);
    }
}

static QDView nd_view(const QDSystems &s)
{
    QDView V{};
    V.nx = s.nx; V.ny = s.ny; V.nb = std::max(s.nx, s.ny);
    V.nc = (size_t)s.nx * s.ny; V.nv = (size_t)(s.nx + 1) * s.ny; V.nh = (size_t)s.nx * (s.ny + 1);
    V.nrow = s.nrow; V.coef = s.coef; V.bnd = s.bnd; V.cs = nullptr;
    return V;
}

// The level table of a chunk of ns systems (front and right-hand-side blocks of each
// level in P.F, P.f), uploaded for the kernels that walk the tree.
static std::vector<NDLevDev> nd_levels(Ctx &c, NDPlan &P, int ns)
{
    const int H = (int)P.lev.size() - 1;
    std::vector<NDLevDev> tab(H + 1);
    for (int l = 0; l <= H; ++l) {
        const NDLevel &L = P.lev[l];
        tab[l] = NDLevDev{L.nn, L.nIp, L.nBp, L.nF, P.F + L.foff * ns, P.f + L.voff * ns, L.up};
    }
    TRT_CHECK(hipMemcpyAsync(P.levdev, tab.data(), tab.size() * sizeof(NDLevDev), hipMemcpyHostToDevice, c.stream));
    return tab;
}

// Normwise backward error of every system of the chunk at x, in the rows scaled by P.W,
// with the residual b - A x left in P.R; true when all are at most tol (and finite).
static bool nd_bkwd(Ctx &c, NDPlan &P, const QDView &V, const double *x, int ns, double tol, double *worst)
{
    hipLaunchKernelGGL(nd_residual_bkwd, dim3(ND_RB_BLOCKS, ns), dim3(256), 0, c.stream, V, P.sysmap, x, P.W, P.R,
                       P.part);
    hipLaunchKernelGGL(nd_bkwd_fin, dim3((ns + 63) / 64), dim3(64), 0, c.stream, P.part, ND_RB_BLOCKS, ns, P.bkwd);
    std::vector<double> hb(ns);
    TRT_CHECK(hipMemcpyAsync(hb.data(), P.bkwd, ns * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    TRT_CHECK(hipStreamSynchronize(c.stream));
    bool ok = true;
    *worst = 0.0;
    for (double b : hb) {
        if (!(b <= tol)) ok = false;
        *worst = std::isfinite(b) ? std::fmax(*worst, b) : INFINITY;
    }
    return ok;
}

// One solve with the stored factors: D = (W A_f)^-1 W R for the chunk's systems, by
// elimination through the fronts bottom-up and back substitution top-down, then x += D.
static void nd_apply(Ctx &c, NDPlan &P, const std::vector<NDLevDev> &tab, int ns, size_t nrow, double *x)
{
    const int H = (int)P.lev.size() - 1;
    const double one = 1.0, mone = -1.0;
    hipLaunchKernelGGL(nd_zero, dim3(nblocks(P.vsz * ns)), dim3(256), 0, c.stream, P.f, P.vsz * ns);
    if (P.ginv) {
        // one launch per level and direction: the right-hand side, the children's parts and
        // the forward elimination; x_B, the back substitution and x_I into D
        for (int l = 0; l <= H; ++l) {
            const NDLevel &L = P.lev[l];
            const int nb = L.nn * ns, bs = L.nF <= 64 ? 64 : 256;
            const NDFwdIn in{P.R, P.W, nrow, L.rows, L.kids, P.levdev, l, ns};
            if (nd_small(L))
                hipLaunchKernelGGL(nd_fwd_lu_small, dim3(nb), dim3(64), 0, c.stream, tab[l].F, tab[l].f,
                                   P.perm + L.poff * ns, L.nIp, L.nBp, L.nF, in);
            else
                hipLaunchKernelGGL(nd_fwd_inv, dim3(nb), dim3(bs), 2 * L.nIp * sizeof(double), c.stream, tab[l].F,
                                   tab[l].f, L.nIp, L.nBp, L.nF, in);
        }
        for (int l = H; l >= 0; --l) {
            const NDLevel &L = P.lev[l];
            const int nb = L.nn * ns, bs = L.nF <= 64 ? 64 : 256;
            hipLaunchKernelGGL(nd_bwd_inv, dim3(nb), dim3(bs), std::max(L.nBp, 1) * sizeof(double), c.stream,
                               tab[l].F, tab[l].f, L.nIp, L.nBp, L.nF, P.D, nrow, ns, L.rows);
        }
        hipLaunchKernelGGL(nd_add, dim3(nblocks((size_t)ns * nrow)), dim3(256), 0, c.stream, x, P.D, P.sysmap, ns,
                           nrow);
        TRT_CHECK(hipGetLastError());
        return;
    }
    for (int l = 0; l <= H; ++l) {
        const NDLevel &L = P.lev[l];
        double *F = tab[l].F, *f = tab[l].f;
        const int nb = L.nn * ns;
        hipLaunchKernelGGL(nd_rhs, dim3(nb), dim3(128), 0, c.stream, P.R, P.W, nrow, ns, L.nIp, L.nF, L.rows, f);
        if (l > 0) hipLaunchKernelGGL(nd_pull_vec, dim3(nb), dim3(128), 0, c.stream, P.levdev, l, ns, L.kids);
        const rocblas_stride sF = (rocblas_stride)L.nF * L.nF, sf = L.nF, sP = L.nIp;
        rocblas_int *piv = P.ipiv + L.poff * ns;
        ND_RB(rocsolver_dgetrs_strided_batched(P.h, rocblas_operation_none, L.nIp, 1, F, L.nF, sF, piv, sP, f, L.nF,
                                               sf, nb));
        if (L.nBp > 0)
            ND_RB(rocblas_dgemm_strided_batched(P.h, rocblas_operation_none, rocblas_operation_none, L.nBp, 1, L.nIp,
                                                &mone, F + L.nIp, L.nF, sF, f, L.nF, sf, &one, f + L.nIp, L.nF, sf,
                                                nb));
    }
    for (int l = H; l >= 0; --l) {
        const NDLevel &L = P.lev[l];
        double *F = tab[l].F, *f = tab[l].f;
        const int nb = L.nn * ns;
        if (L.nBp > 0) {
            hipLaunchKernelGGL(nd_gather, dim3(nb), dim3(128), 0, c.stream, P.D, nrow, P.ident, ns, L.nIp, L.nBp,
                               L.nF, L.rows, f);
            const rocblas_stride sF = (rocblas_stride)L.nF * L.nF, sf = L.nF;
            ND_RB(rocblas_dgemm_strided_batched(P.h, rocblas_operation_none, rocblas_operation_none, L.nIp, 1, L.nBp,
                                                &mone, F + (size_t)L.nIp * L.nF, L.nF, sF, f + L.nIp, L.nF, sf, &one,
                                                f, L.nF, sf, nb));
        }
        hipLaunchKernelGGL(nd_scatter, dim3(nb), dim3(128), 0, c.stream, P.D, nrow, P.ident, ns, L.nIp, L.nF, L.rows,
                           f);
    }
    hipLaunchKernelGGL(nd_add, dim3(nblocks((size_t)ns * nrow)), dim3(256), 0, c.stream, x, P.D, P.sysmap, ns, nrow);
    TRT_CHECK(hipGetLastError());
}

void nd_destroy(Ctx &c)
{
    for (QDSystems *s : {&c.mg, &c.grey}) {
        if (!s->nd) continue;
        if (s->nd->h) rocblas_destroy_handle(s->nd->h);
        delete s->nd;     // device buffers belong to the context (dalloc)
        s->nd = nullptr;
    }
}

static NDPlan &nd_plan_of(Ctx &c, QDSystems &s)
{
    if (s.nd && (s.nd->nx != s.nx || s.nd->ny != s.ny || s.nd->nrow != s.nrow)) {
        if (s.nd->h) rocblas_destroy_handle(s.nd->h);
        delete s.nd;
        s.nd = nullptr;
    }
    if (!s.nd) s.nd = nd_plan(c, s);
    ND_RB(rocblas_set_stream(s.nd->h, c.stream));
    return *s.nd;
}

// Solve the listed systems of s; their solutions replace s.x.  When they fit one
// chunk their factors stay in the plan, for qd_nd_reuse.
void qd_nd(Ctx &c, QDSystems &s, const std::vector<int> &which)
{
    if (which.empty()) return;
    NDPlan &P = nd_plan_of(c, s);
    P.fact_ok = false;
    P.ginv = false;
    const QDView V = nd_view(s);
    const int H = (int)P.lev.size() - 1;
    const double one = 1.0, mone = -1.0;
    for (size_t c0 = 0; c0 < which.size(); c0 += P.cap) {
        const int ns = (int)std::min<size_t>(P.cap, which.size() - c0);
        TRT_CHECK(hipMemcpyAsync(P.sysmap, which.data() + c0, ns * sizeof(int), hipMemcpyHostToDevice, c.stream));
        const std::vector<NDLevDev> tab = nd_levels(c, P, ns);
        hipLaunchKernelGGL(nd_zero, dim3(nblocks(P.fsz * ns)), dim3(256), 0, c.stream, P.F, P.fsz * ns);
        hipLaunchKernelGGL(nd_zero, dim3(nblocks(P.vsz * ns)), dim3(256), 0, c.stream, P.f, P.vsz * ns);
        hipLaunchKernelGGL(nd_rowscale, dim3(nblocks((size_t)ns * s.nrow)), dim3(256), 0, c.stream, V, P.sysmap, ns,
                           P.W);
        size_t ioff = 0;
        for (int l = 0; l <= H; ++l) {
            const NDLevel &L = P.lev[l];
            double *F = tab[l].F, *f = tab[l].f;
            const int nb = L.nn * ns;
            hipLaunchKernelGGL(nd_assemble, dim3(nb), dim3(128), 0, c.stream, V, P.sysmap, ns, L.nIp, L.nF, L.rows,
                               L.terms, P.W, F, f);
            if (l > 0) {
                // children's border blocks are at most the largest of the levels below
                int mb = 0;
                for (int q = 0; q < l; ++q) mb = std::max(mb, P.lev[q].nBp);
                const int chunks = (int)std::min<long long>(64, std::max<long long>(1, ((long long)mb * mb + 1023) / 1024));
                for (int k = 0; k < 2; ++k)
                    hipLaunchKernelGGL(nd_pull_child, dim3(nb, chunks), dim3(256), 0, c.stream, P.levdev, l, ns,
                                       L.kids, k);
            }
            TRT_CHECK(hipGetLastError());
            const rocblas_stride sF = (rocblas_stride)L.nF * L.nF, sf = L.nF, sP = L.nIp;
            rocblas_int *piv = P.ipiv + L.poff * ns;
            if (nd_small(L)) {
                // small fronts (the leaves and the levels above them): one kernel, one wavefront each
                double *Fa = F, *fa = f;
                int nFa = L.nF, nIpa = L.nIp, nba = nb;
                size_t sFa = (size_t)sF, sfa = (size_t)sf;
                rocblas_int *pa = piv, *ia = P.info + ioff;
                int *perma = P.perm + L.poff * ns;
                void *args[] = {&Fa, &fa, &nFa, &nIpa, &sFa, &sfa, &pa, &ia, &perma, &nba};
                TRT_CHECK(hipLaunchKernel(nd_front_small_kernel(L.nIp), dim3(nb), dim3(64), args, 0, c.stream));
                ioff += (size_t)nb;
                continue;
            }
            ND_RB(rocsolver_dgetrf_strided_batched(P.h, L.nIp, L.nIp, F, L.nF, sF, piv, sP, P.info + ioff, nb));
            ioff += (size_t)nb;
            ND_RB(rocsolver_dgetrs_strided_batched(P.h, rocblas_operation_none, L.nIp, 1, F, L.nF, sF, piv, sP, f,
                                                   L.nF, sf, nb));
            if (L.nBp > 0) {
                double *FIB = F + (size_t)L.nIp * L.nF, *FBI = F + L.nIp, *FBB = F + L.nIp + (size_t)L.nIp * L.nF;
                ND_RB(rocsolver_dgetrs_strided_batched(P.h, rocblas_operation_none, L.nIp, L.nBp, F, L.nF, sF, piv,
                                                       sP, FIB, L.nF, sF, nb));
                ND_RB(rocblas_dgemm_strided_batched(P.h, rocblas_operation_none, rocblas_operation_none, L.nBp, L.nBp,
                                                    L.nIp, &mone, FBI, L.nF, sF, FIB, L.nF, sF, &one, FBB, L.nF, sF,
                                                    nb));
                ND_RB(rocblas_dgemm_strided_batched(P.h, rocblas_operation_none, rocblas_operation_none, L.nBp, 1,
                                                    L.nIp, &mone, FBI, L.nF, sF, f, L.nF, sf, &one, f + L.nIp, L.nF,
                                                    sf, nb));
            }
        }
        hipLaunchKernelGGL(nd_info_max, dim3(1), dim3(1024), 0, c.stream, P.info, ioff, P.imax);
        // back substitution, top-down
        for (int l = H; l >= 0; --l) {
            const NDLevel &L = P.lev[l];
            double *F = tab[l].F, *f = tab[l].f;
            const int nb = L.nn * ns;
            if (L.nBp > 0) {
                hipLaunchKernelGGL(nd_gather, dim3(nb), dim3(128), 0, c.stream, s.x, s.nrow, P.sysmap, ns, L.nIp,
                                   L.nBp, L.nF, L.rows, f);
                const rocblas_stride sF = (rocblas_stride)L.nF * L.nF, sf = L.nF;
                ND_RB(rocblas_dgemm_strided_batched(P.h, rocblas_operation_none, rocblas_operation_none, L.nIp, 1,
                                                    L.nBp, &mone, F + (size_t)L.nIp * L.nF, L.nF, sF, f + L.nIp,
                                                    L.nF, sf, &one, f, L.nF, sf, nb));
            }
            hipLaunchKernelGGL(nd_scatter, dim3(nb), dim3(128), 0, c.stream, s.x, s.nrow, P.sysmap, ns, L.nIp, L.nF,
                               L.rows, f);
        }
        TRT_CHECK(hipGetLastError());
        int imax = 0;
        TRT_CHECK(hipMemcpyAsync(&imax, P.imax, sizeof(int), hipMemcpyDeviceToHost, c.stream));
        TRT_CHECK(hipStreamSynchronize(c.stream));
        if (imax > 0) throw Error("nested dissection: a front's pivot block is singular");
        // factor-reuse mode: the pivot blocks' LU factors become their inverses, so that
        // later solves with these factors are matrix-vector products (bandwidth-bound)
        // instead of batched triangular solves
        if (c.qd_solver == 0) {
            for (int l = 0; l <= H; ++l) {
                const NDLevel &L = P.lev[l];
                if (nd_small(L)) continue;       // small fronts keep their LU (nd_fwd_lu_small)
                const rocblas_stride sF = (rocblas_stride)L.nF * L.nF, sP = L.nIp;
                ND_RB(rocsolver_dgetri_strided_batched(P.h, L.nIp, tab[l].F, L.nF, sF, P.ipiv + L.poff * ns, sP,
                                                       P.info, L.nn * ns));
            }
            P.ginv = true;
        }
        // iterative refinement with the stored factors while the worst backward
        // error of the chunk is above 1e-14 (at most three steps): pivoting inside
        // each front only can leave it above rounding
        for (int it = 0;; ++it) {
            double worst = 0.0;
            const bool ok = nd_bkwd(c, P, V, s.x, ns, 1e-14, &worst);
            if (c.verbose > 2)
                fprintf(stderr, "[qd_nd] chunk of %d: backward error %.3e%s after %d refinements\n", ns, worst,
                        ok ? "" : " (or not finite)", it);
            if (ok || it == 3) break;
            nd_apply(c, P, tab, ns, s.nrow, s.x);
        }
        if (c0 == 0 && (size_t)ns == which.size()) {
            P.fact_ok = true;
            P.fact_which = which;
        }
    }
    if (c.verbose > 1)
        fprintf(stderr, "[qd_nd] %zu systems, %zu levels, %.2e flop per system\n", which.size(), P.lev.size(), P.flops);
}

// The listed systems solved with the factors that the last numeric phase of s left in
// its plan, for the same systems, whose matrices have changed since (the Newton and
// MLOQD iterations of a time step change them a little): iterative refinement from the
// initial guess in s.x, x += (W A_f)^-1 W (b - A x), until every normwise backward
// error is at most 1e-14, the acceptance of qd_nd.  Returns the refinement steps taken,
// or -1 -- s.x then undefined, the caller solves afresh -- when the plan holds no
// factors of these systems, or when the observed contraction (the last step's ratio of
// worst backward errors) is above 1/2 or predicts more than max_steps steps in all:
// refinement stops as soon as it would cost more than it saves.
void qd_nd_prepare(Ctx &c, QDSystems &s)
{
    if (s.nsys > 0) (void)nd_plan_of(c, s);
}

void qd_nd_forget(QDSystems &s)
{
    if (s.nd) s.nd->fact_ok = false;
}

int qd_nd_reuse(Ctx &c, QDSystems &s, const std::vector<int> &which, int max_steps)
{
    if (which.empty() || !s.nd || !s.nd->fact_ok || s.nd->fact_which != which) return -1;
    NDPlan &P = nd_plan_of(c, s);
    const int ns = (int)which.size();
    const QDView V = nd_view(s);
    TRT_CHECK(hipMemcpyAsync(P.sysmap, which.data(), ns * sizeof(int), hipMemcpyHostToDevice, c.stream));
    const std::vector<NDLevDev> tab = nd_levels(c, P, ns);
    double prev = INFINITY;
    for (int it = 0;; ++it) {
        double worst = 0.0;
        const bool ok = nd_bkwd(c, P, V, s.x, ns, 1e-14, &worst);
        if (c.verbose > 2)
            fprintf(stderr, "[qd_nd reuse] %d systems: backward error %.3e after %d steps\n", ns, worst, it);
        if (ok) return it;
        if (it == max_steps || !(worst <= 0.5 * prev)) return -1;
        if (it > 0) {
            const double need = it + std::log(1e-14 / worst) / std::log(worst / prev);
            if (!(need <= max_steps)) return -1;
        }
        prev = worst;
        nd_apply(c, P, tab, ns, s.nrow, s.x);
    }
}

}  // namespace trt
