// Direct solve of reduced QD systems: block-tridiagonal LU over grid rows.
//
// QD_FV falls back to a direct solve when its Krylov solve fails
// (MainSrc/QD_SOLVERS.f08:492-500, dense Gaussian elimination).  A dense N x N
// matrix is out of reach at N = 197,120, but the reduced system is block
// tridiagonal when its unknowns are grouped by grid row:
//   block j < N_y : [E_avg(0..N_x-1, j) | E_edgV(0..N_x, j) | E_edgH(0..N_x-1, j)]  (m = 3 N_x + 1)
//   block N_y     : [E_edgH(0..N_x-1, N_y)], padded to m with identity rows.
// The energy balance and x-momentum rows of grid row j couple blocks j and j+1
// (through E_edgH(., j+1)); the y-momentum rows of the faces below row j couple
// blocks j-1 and j.  Block LU with partial pivoting inside each diagonal block:
//   D_j = A_jj - L_j S_{j-1},  S_j = D_j^{-1} U_j,  y_j = D_j^{-1} (b_j - L_j y_{j-1}),
//   x_j = y_j - S_j x_{j+1},
// on dense m x m blocks with rocSOLVER getrf/getrs and rocBLAS gemm, batched over
// the systems being solved (strided batched calls).
#include "qd_view.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cstdio>

namespace trt {

#define RB_CHECK(expr)                                                                     \
    do {                                                                                   \
        rocblas_status _s = (expr);                                                        \
        if (_s != rocblas_status_success)                                                  \
            throw Error(std::string(#expr) + ": rocBLAS status " + std::to_string((int)_s)); \
    } while (0)

struct DirectWork {
    rocblas_handle h = nullptr;
    int m = 0, nb = 0, cap = 0;          // block size, blocks, systems per chunk
    double *D = nullptr, *L = nullptr, *U = nullptr, *Y = nullptr;
    rocblas_int *ipiv = nullptr, *info = nullptr;
    int *sys = nullptr;                  // systems of the chunk
};

namespace {

__device__ __forceinline__ void block_of(const QDView &V, int kind, int i, int j, int &blk, int &loc)
{
    const int nx = V.nx;
    if (kind == 0) { blk = j; loc = i; }
    else if (kind == 1) { blk = j; loc = nx + i; }
    else if (j < V.ny) { blk = j; loc = 2 * nx + 1 + i; }
    else { blk = V.ny; loc = i; }
}

// scatter the rows of the chunk's systems into the dense blocks (column-major m x m)
__global__ void assemble_blocks(QDView V, const int *sys, int nsys, int m, int nb, double *D, double *L, double *U,
                                double *Y)
{
    const size_t per = (size_t)nb * m;
    const size_t tot = per * nsys;
    const size_t bs = (size_t)m * m;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < tot; t += (size_t)gridDim.x * blockDim.x) {
        const int s = (int)(t / per);
        const size_t r = t - (size_t)s * per;
        const int jr = (int)(r / m), lr = (int)(r - (size_t)jr * m);
        const size_t base = (size_t)s * nb * bs + (size_t)jr * bs;
        double *Dj = D + base, *Lj = L + base, *Uj = U + base;
        double *yj = Y + (size_t)s * per + (size_t)jr * m;
        // which unknown is local row lr of block jr
        int kind, i, j;
        if (jr < V.ny) {
            j = jr;
            if (lr < V.nx) { kind = 0; i = lr; }
            else if (lr < 2 * V.nx + 1) { kind = 1; i = lr - V.nx; }
            else { kind = 2; i = lr - 2 * V.nx - 1; }
        } else {
            j = V.ny; kind = 2; i = lr;
            if (lr >= V.nx) {                // padding row
                Dj[(size_t)lr * m + lr] = 1.0;
                yj[lr] = 0.0;
                continue;
            }
        }
        const size_t row = qd_index(V, kind, i, j);
        double diag, rhs;
        qd_row_visit(V, sys[s], row, [&](int k2, int i2, int j2, double a) {
            int bc, lc;
            block_of(V, k2, i2, j2, bc, lc);
            if (bc == jr) Dj[(size_t)lc * m + lr] += a;
            else if (bc == jr - 1) Lj[(size_t)lc * m + lr] += a;
            else Uj[(size_t)lc * m + lr] += a;
        }, diag, rhs);
        yj[lr] = rhs;
    }
}

// x[sys] (unknown layout) <- Y (block layout)
__global__ void scatter_solution(QDView V, const int *sys, int nsys, int m, int nb, const double *Y, double *x)
{
    const size_t per = (size_t)nb * m;
    const size_t tot = per * nsys;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < tot; t += (size_t)gridDim.x * blockDim.x) {
        const int s = (int)(t / per);
        const size_t r = t - (size_t)s * per;
        const int jr = (int)(r / m), lr = (int)(r - (size_t)jr * m);
        int kind, i, j;
        if (jr < V.ny) {
            j = jr;
            if (lr < V.nx) { kind = 0; i = lr; }
            else if (lr < 2 * V.nx + 1) { kind = 1; i = lr - V.nx; }
            else { kind = 2; i = lr - 2 * V.nx - 1; }
        } else {
            if (lr >= V.nx) continue;
            j = V.ny; kind = 2; i = lr;
        }
        x[(size_t)sys[s] * V.nrow + qd_index(V, kind, i, j)] = Y[t];
    }
}

__global__ void zero_kernel(double *p, size_t n)
{
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) p[k] = 0.0;
}

int nblocks_for(size_t n) { return (int)std::min<size_t>(std::max<size_t>((n + 255) / 256, 1), 65536); }

}  // namespace

void direct_destroy(Ctx &c)
{
    if (!c.direct) return;
    if (c.direct->h) rocblas_destroy_handle(c.direct->h);
    delete c.direct;
    c.direct = nullptr;
}

// Solve the systems listed in which[] of s by block LU; their solutions replace s.x.
void qd_direct(Ctx &c, QDSystems &s, const std::vector<int> &which)
{
    if (which.empty()) return;
    const int nx = s.nx, ny = s.ny;
    const int m = 3 * nx + 1, nb = ny + 1;
    const size_t bs = (size_t)m * m;
    const size_t per_sys = 3 * (size_t)nb * bs * sizeof(double);
    if (!c.direct) {
        c.direct = new DirectWork();
        RB_CHECK(rocblas_create_handle(&c.direct->h));
    }
    DirectWork &w = *c.direct;
    RB_CHECK(rocblas_set_stream(w.h, c.stream));
    if (w.m != m || w.nb != nb) {
        const double budget = 24.0 * (1u << 30);          // bytes of dense blocks per chunk
        int cap = (int)std::max<double>(1.0, std::floor(budget / (double)per_sys));
        cap = std::min(cap, s.nsys);
        w.m = m; w.nb = nb; w.cap = cap;
        w.D = static_cast<double *>(c.dalloc((size_t)cap * nb * bs * 8));
        w.L = static_cast<double *>(c.dalloc((size_t)cap * nb * bs * 8));
        w.U = static_cast<double *>(c.dalloc((size_t)cap * nb * bs * 8));
        w.Y = static_cast<double *>(c.dalloc((size_t)cap * nb * m * 8));
        w.ipiv = static_cast<rocblas_int *>(c.dalloc((size_t)cap * nb * m * sizeof(rocblas_int)));
        w.info = static_cast<rocblas_int *>(c.dalloc((size_t)cap * nb * sizeof(rocblas_int)));
        w.sys = static_cast<int *>(c.dalloc((size_t)std::max(cap, 1) * sizeof(int)));
    }
    QDView V{};
    V.nx = nx; V.ny = ny; V.nb = std::max(nx, ny);
    V.nc = (size_t)nx * ny; V.nv = (size_t)(nx + 1) * ny; V.nh = (size_t)nx * (ny + 1);
    V.nrow = s.nrow; V.coef = s.coef; V.bnd = s.bnd; V.cs = nullptr;
    const rocblas_stride sD = (rocblas_stride)((size_t)nb * bs), sY = (rocblas_stride)((size_t)nb * m);
    const rocblas_stride sP = (rocblas_stride)((size_t)nb * m);
    const double one = 1.0, mone = -1.0;
    for (size_t c0 = 0; c0 < which.size(); c0 += w.cap) {
        const int ns = (int)std::min<size_t>(w.cap, which.size() - c0);
        TRT_CHECK(hipMemcpyAsync(w.sys, which.data() + c0, ns * sizeof(int), hipMemcpyHostToDevice, c.stream));
        const size_t nmat = (size_t)ns * nb * bs;
        hipLaunchKernelGGL(zero_kernel, dim3(nblocks_for(nmat)), dim3(256), 0, c.stream, w.D, nmat);
        hipLaunchKernelGGL(zero_kernel, dim3(nblocks_for(nmat)), dim3(256), 0, c.stream, w.L, nmat);
        hipLaunchKernelGGL(zero_kernel, dim3(nblocks_for(nmat)), dim3(256), 0, c.stream, w.U, nmat);
        const size_t nrows = (size_t)ns * nb * m;
        hipLaunchKernelGGL(assemble_blocks, dim3(nblocks_for(nrows)), dim3(256), 0, c.stream, V, w.sys, ns, m, nb, w.D,
                           w.L, w.U, w.Y);
        TRT_CHECK(hipGetLastError());
        for (int j = 0; j < nb; ++j) {
            double *Dj = w.D + (size_t)j * bs, *Lj = w.L + (size_t)j * bs, *Uj = w.U + (size_t)j * bs;
            double *yj = w.Y + (size_t)j * m;
            if (j > 0) {
                const double *Sp = w.U + (size_t)(j - 1) * bs;
                const double *yp = w.Y + (size_t)(j - 1) * m;
                RB_CHECK(rocblas_dgemm_strided_batched(w.h, rocblas_operation_none, rocblas_operation_none, m, m, m,
                                                       &mone, Lj, m, sD, Sp, m, sD, &one, Dj, m, sD, ns));
                RB_CHECK(rocblas_dgemm_strided_batched(w.h, rocblas_operation_none, rocblas_operation_none, m, 1, m,
                                                       &mone, Lj, m, sD, yp, m, sY, &one, yj, m, sY, ns));
            }
            rocblas_int *pj = w.ipiv + (size_t)j * m;
            RB_CHECK(rocsolver_dgetrf_strided_batched(w.h, m, m, Dj, m, sD, pj, sP, w.info + (size_t)j * w.cap, ns));
            if (j < nb - 1)
                RB_CHECK(rocsolver_dgetrs_strided_batched(w.h, rocblas_operation_none, m, m, Dj, m, sD, pj, sP, Uj, m,
                                                          sD, ns));
            RB_CHECK(rocsolver_dgetrs_strided_batched(w.h, rocblas_operation_none, m, 1, Dj, m, sD, pj, sP, yj, m, sY,
                                                      ns));
        }
        for (int j = nb - 2; j >= 0; --j) {
            const double *Sj = w.U + (size_t)j * bs;
            RB_CHECK(rocblas_dgemm_strided_batched(w.h, rocblas_operation_none, rocblas_operation_none, m, 1, m, &mone,
                                                   Sj, m, sD, w.Y + (size_t)(j + 1) * m, m, sY, &one,
                                                   w.Y + (size_t)j * m, m, sY, ns));
        }
        hipLaunchKernelGGL(scatter_solution, dim3(nblocks_for(nrows)), dim3(256), 0, c.stream, V, w.sys, ns, m, nb, w.Y,
                           s.x);
        TRT_CHECK(hipGetLastError());
    }
    if (c.verbose) fprintf(stderr, "[qd_direct] %zu systems by block LU (m = %d, %d blocks)\n", which.size(), m, nb);
}

}  // namespace trt
