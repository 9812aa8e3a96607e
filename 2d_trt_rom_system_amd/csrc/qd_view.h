// Matrix-free view of the reduced QD system (QD_FV, MainSrc/QD_SOLVERS.f08:40-538)
// shared by the Krylov solver (loqd.hip) and the direct fallback (qd_direct.hip).
#pragma once

#include "trt_internal.h"

namespace trt {

enum { EBL, EBB, EBC, EBR, EBT, MXC, MXR, MXB, MXT, MYC, MYT, MYL, MYR, ERHS, MXRHS, MYRHS, NCOEF };
enum { CPL, CPB, CPR, CPT, BCL, BCB, BCR, BCT, NBND };

// ---------------------------------------------------------------------------
// matrix-free row of the reduced QD system
// ---------------------------------------------------------------------------
struct QDView {
    int nx, ny, nb;
    size_t nc, nv, nh, nrow;
    const double *coef;   // [sys][NCOEF][nc]
    const double *bnd;    // [sys][NBND][nb]
    const double *cs;     // [sys][nrow] column scales: the row computes (A S y)_row for x = S y
};

// The equation owned by unknown `row` of system `sys`: term(kind, i, j, a) is called
// for every coefficient a multiplying unknown (kind, i, j) (kind 0 = E_avg(i,j),
// 1 = E_edgV(i,j), 2 = E_edgH(i,j)); diag and rhs receive A_row,row and b_row.
// (host-callable too: the direct solver's plan walks the stencil on the host)
template <typename F>
__host__ __device__ __forceinline__ void qd_row_visit(const QDView &V, int sys, size_t row, F term, double &diag,
                                                      double &rhs)
{
    const int nx = V.nx, ny = V.ny;
    const double *cf = V.coef + (size_t)sys * NCOEF * V.nc;
    const double *bd = V.bnd + (size_t)sys * NBND * V.nb;
#define CF(k, i, j) cf[(size_t)(k) * V.nc + (size_t)(i) + (size_t)nx * (j)]
    if (row < V.nc) {                       // energy balance of cell (i,j)
        const int i = (int)(row % nx), j = (int)(row / nx);
        term(1, i, j, CF(EBL, i, j)); term(2, i, j, CF(EBB, i, j)); term(0, i, j, CF(EBC, i, j));
        term(1, i + 1, j, CF(EBR, i, j)); term(2, i, j + 1, CF(EBT, i, j));
        diag = CF(EBC, i, j);
        rhs = CF(ERHS, i, j);
    } else if (row < V.nc + V.nv) {         // vertical face f of row j
        const size_t r = row - V.nc;
        const int f = (int)(r % (nx + 1)), j = (int)(r / (nx + 1));
        if (f == 0) {                       // left boundary condition (QS:92-97, 262-267)
            term(1, 0, j, bd[CPL * V.nb + j]); term(2, 0, j, -CF(MXB, 0, j)); term(0, 0, j, CF(MXC, 0, j));
            term(2, 0, j + 1, -CF(MXT, 0, j));
            diag = bd[CPL * V.nb + j];
            rhs = bd[BCL * V.nb + j];
        } else if (f == nx) {               // right boundary condition (QS:198-203, 364-369)
            const int i = nx - 1;
            term(2, i, j, -CF(MXB, i, j)); term(0, i, j, -CF(MXC, i, j)); term(1, nx, j, bd[CPR * V.nb + j]);
            term(2, i, j + 1, -CF(MXT, i, j));
            diag = bd[CPR * V.nb + j];
            rhs = bd[BCR * V.nb + j];
        } else {                            // x-momentum between cells f-1 and f (QS:117-125)
            const int i = f - 1;
            term(2, i, j, CF(MXB, i, j)); term(0, i, j, CF(MXC, i, j)); term(1, f, j, CF(MXR, i, j));
            term(2, i, j + 1, CF(MXT, i, j)); term(2, f, j, -CF(MXB, f, j)); term(0, f, j, CF(MXC, f, j));
            term(2, f, j + 1, -CF(MXT, f, j));
            diag = CF(MXR, i, j);
            rhs = CF(MXRHS, i, j);
        }
    } else {                                // horizontal face f of column i
        const size_t r = row - V.nc - V.nv;
        const int i = (int)(r % nx), f = (int)(r / nx);
        if (f == 0) {                       // bottom boundary condition (QS:100-105, 168-173)
            term(1, i, 0, -CF(MYL, i, 0)); term(2, i, 0, bd[CPB * V.nb + i]); term(0, i, 0, CF(MYC, i, 0));
            term(1, i + 1, 0, -CF(MYR, i, 0));
            diag = bd[CPB * V.nb + i];
            rhs = bd[BCB * V.nb + i];
        } else if (f == ny) {               // top boundary condition (QS:303-309, 388-393)
            const int j = ny - 1;
            term(1, i, j, -CF(MYL, i, j)); term(0, i, j, -CF(MYC, i, j)); term(1, i + 1, j, -CF(MYR, i, j));
            term(2, i, ny, bd[CPT * V.nb + i]);
            diag = bd[CPT * V.nb + i];
            rhs = bd[BCT * V.nb + i];
        } else {                            // y-momentum between cells f-1 and f (QS:128-137)
            const int j = f - 1;
            term(1, i, j, CF(MYL, i, j)); term(0, i, j, CF(MYC, i, j)); term(1, i + 1, j, CF(MYR, i, j));
            term(2, i, f, CF(MYT, i, j)); term(1, i, f, -CF(MYL, i, f)); term(0, i, f, CF(MYC, i, f));
            term(1, i + 1, f, -CF(MYR, i, f));
            diag = CF(MYT, i, j);
            rhs = CF(MYRHS, i, j);
        }
    }
#undef CF
}

__host__ __device__ __forceinline__ size_t qd_index(const QDView &V, int kind, int i, int j)
{
    return kind == 0 ? (size_t)i + (size_t)V.nx * j
         : kind == 1 ? V.nc + (size_t)i + (size_t)(V.nx + 1) * j
                     : V.nc + V.nv + (size_t)i + (size_t)V.nx * j;
}

}  // namespace trt
