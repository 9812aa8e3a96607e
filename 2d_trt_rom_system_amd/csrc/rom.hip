// POD reconstruction of a QD-factor field for the ROM runs' step: POD_RECONSTRUCT
// (MainSrc/POD_ROUTINES.f08:372-432) and LINEAR_INTERPOLATION between two data
// times (MainSrc/GRID_FUNCTIONS.f08:819-836), as POD_RECONSTRUCT_fg calls them
// (POD_ROUTINES.f08:28-130) for data on the simulation grid.
//
// A basis (centering vector C, singular values S, left vectors U, right vectors
// V) lives on the device; a reconstruction writes one context array (normally a
// *_next array of trt_step_given_qdf):
//   per group (PODgsum 0):  f[g][i] = sum_r S_g[r] U_g[r][i] V_g[t][r] + C[g][i]
//   all groups (PODgsum 1): f[i]    = sum_r S[r]   U[r][i]   V[t][r]   + C[i]
// with the reference's storage: U group blocks of rank(g) columns of glen values,
// S and V concatenated over the groups, V of a group t-major (rank(g) values per
// data time).  Between two data times each element is y1 + (x - x1)(y2 - y1)/(x2 - x1).
// One pass over U: a thread per element, both time points accumulated together.
#include "trt_internal.h"

#include <cmath>
#include <cstdint>
#include <numeric>

struct trt_pod {
    trt_ctx *ctx = nullptr;
    std::string name;
    int gsum = 0, ng = 0, nt = 0;
    size_t len = 0, glen = 0;
    double *C = nullptr, *S = nullptr, *U = nullptr, *V = nullptr;
    long long *uoff = nullptr;   // per group: offset of its U block
    int *soff = nullptr;         // per group: offset of its S values (p3)
    int *rank = nullptr;
    size_t ubytes = 0;
};

// DMD reconstruction (DMD_RECONSTRUCT, MainSrc/DMD_ROUTINES.f08:28-103): the real part of
// sum_r B_r W_r[i] exp(L_r t), plus C[i]; W in U's block layout, complex (re, im) pairs.
// The factors exp(L_r t) are formed on the host per call (rank values).
struct trt_dmd {
    trt_ctx *ctx = nullptr;
    std::string name;
    int gsum = 0, ng = 0;
    size_t len = 0, glen = 0;
    std::vector<double> L, B;                // host: (re, im) pairs, sum of ranks
    std::vector<int> rank_h;
    double *C = nullptr, *W = nullptr, *B_dev = nullptr, *E = nullptr;
    long long *woff = nullptr;
    int *soff = nullptr, *rank = nullptr;
};

namespace {

// BCg_next -> Cg_L/B/R/T_next: per group the data vector holds the left side (N_y
// values), the bottom (N_x), the right (N_y), the top (N_x) (POD_ROUTINES.f08:279-352,
// with the data grid the simulation grid)
__global__ void bcg_split_kernel(const double *__restrict__ bc, double *__restrict__ L, double *__restrict__ B,
                                 double *__restrict__ R, double *__restrict__ T, int nx, int ny, int ng)
{
    const int per = 2 * (nx + ny);
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < per * ng; k += gridDim.x * blockDim.x) {
        const int g = k / per, j = k - g * per;
        const double v = bc[k];
        if (j < ny) L[j + ny * g] = v;
        else if (j < ny + nx) B[(j - ny) + nx * g] = v;
        else if (j < 2 * ny + nx) R[(j - ny - nx) + ny * g] = v;
        else T[(j - 2 * ny - nx) + nx * g] = v;
    }
}

__global__ void dmd_kernel(const double *__restrict__ C, const double2 *__restrict__ W, const double2 *__restrict__ B,
                           const double2 *__restrict__ E, const long long *__restrict__ woff,
                           const int *__restrict__ soff, const int *__restrict__ rank, size_t len, size_t glen,
                           int gsum, double *__restrict__ out)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < len; i += (size_t)gridDim.x * blockDim.x) {
        const int g = gsum ? 0 : (int)(i / glen);
        const size_t ii = gsum ? i : i - (size_t)g * glen;
        const size_t stride = gsum ? len : glen;
        const int rk = rank[g], p3 = soff[g];
        const double2 *w = W + woff[g] + ii;
        double acc = 0.0;
#pragma unroll 4
        for (int r = 0; r < rk; ++r) {
            const double2 b = B[p3 + r], e = E[p3 + r], wv = w[(size_t)r * stride];
            // (B W) exp(L t), left to right as the reference multiplies
            const double bwr = b.x * wv.x - b.y * wv.y, bwi = b.x * wv.y + b.y * wv.x;
            acc += bwr * e.x - bwi * e.y;
        }
        out[i] = acc + C[i];
    }
}

__global__ void pod_kernel(const double *__restrict__ C, const double *__restrict__ S, const double *__restrict__ U,
                           const double *__restrict__ V, const long long *__restrict__ uoff,
                           const int *__restrict__ soff, const int *__restrict__ rank, int nt, size_t len, size_t glen,
                           int gsum, int t1, int t2, double x, double x1, double x2, double *__restrict__ out)
{
    const bool interp = t1 != t2;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < len; i += (size_t)gridDim.x * blockDim.x) {
        const int g = gsum ? 0 : (int)(i / glen);
        const size_t ii = gsum ? i : i - (size_t)g * glen;
        const size_t stride = gsum ? len : glen;
        const int rk = rank[g], p3 = soff[g];
        const double *u = U + uoff[g] + ii;
        const double *s = S + p3;
        const double *v1 = V + (gsum ? 0 : (size_t)nt * p3) + (size_t)t1 * rk;
        const double *v2 = V + (gsum ? 0 : (size_t)nt * p3) + (size_t)t2 * rk;
        double a1 = 0.0, a2 = 0.0;
        // unrolled so several U loads are in flight per thread (the sum keeps its order)
#pragma unroll 8
        for (int r = 0; r < rk; ++r) {
            const double su = s[r] * u[(size_t)r * stride];
            a1 += su * v1[r];
            if (interp) a2 += su * v2[r];
        }
        const double y1 = a1 + C[i];
        out[i] = interp ? y1 + (x - x1) * ((a2 + C[i]) - y1) / (x2 - x1) : y1;
    }
}

// Two elements per thread with 16-byte loads, for an even group length: a pair
// never straddles two groups and every pair starts 16-byte aligned (group blocks of
// U start at multiples of glen).  Each element's sum keeps the scalar kernel's order.
__global__ void pod_kernel2(const double *__restrict__ C, const double *__restrict__ S, const double *__restrict__ U,
                            const double *__restrict__ V, const long long *__restrict__ uoff,
                            const int *__restrict__ soff, const int *__restrict__ rank, int nt, size_t len,
                            size_t glen, int gsum, int t1, int t2, double x, double x1, double x2,
                            double *__restrict__ out)
{
    const bool interp = t1 != t2;
    const size_t npair = len / 2;
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < npair; k += (size_t)gridDim.x * blockDim.x) {
        const size_t i = 2 * k;
        const int g = gsum ? 0 : (int)(i / glen);
        const size_t ii = gsum ? i : i - (size_t)g * glen;
        const size_t stride = gsum ? len : glen;
        const int rk = rank[g], p3 = soff[g];
        const double *u = U + uoff[g] + ii;
        const double *s = S + p3;
        const double *v1 = V + (gsum ? 0 : (size_t)nt * p3) + (size_t)t1 * rk;
        const double *v2 = V + (gsum ? 0 : (size_t)nt * p3) + (size_t)t2 * rk;
        double a1x = 0.0, a1y = 0.0, a2x = 0.0, a2y = 0.0;
#pragma unroll 8
        for (int r = 0; r < rk; ++r) {
            const double2 uu = *reinterpret_cast<const double2 *>(u + (size_t)r * stride);
            const double sx = s[r] * uu.x, sy = s[r] * uu.y;
            a1x += sx * v1[r];
            a1y += sy * v1[r];
            if (interp) {
                a2x += sx * v2[r];
                a2y += sy * v2[r];
            }
        }
        const double2 cc = *reinterpret_cast<const double2 *>(C + i);
        const double y1x = a1x + cc.x, y1y = a1y + cc.y;
        double2 o;
        o.x = interp ? y1x + (x - x1) * ((a2x + cc.x) - y1x) / (x2 - x1) : y1x;
        o.y = interp ? y1y + (x - x1) * ((a2y + cc.y) - y1y) / (x2 - x1) : y1y;
        *reinterpret_cast<double2 *>(out + i) = o;
    }
}

void pod_launch(const trt_pod *p, double *dst, hipStream_t st, int t1, int t2, double x, double x1, double x2)
{
    const size_t seg = p->gsum ? p->len : p->glen;
    if (seg % 2 == 0 && (reinterpret_cast<uintptr_t>(dst) % 16) == 0) {
        const int blocks = (int)std::min<size_t>((p->len / 2 + 255) / 256, 8192);
        hipLaunchKernelGGL(pod_kernel2, dim3(blocks), dim3(256), 0, st, p->C, p->S, p->U, p->V, p->uoff, p->soff,
                           p->rank, p->nt, p->len, p->glen, p->gsum, t1, t2, x, x1, x2, dst);
    } else {
        const int blocks = (int)std::min<size_t>((p->len + 255) / 256, 8192);
        hipLaunchKernelGGL(pod_kernel, dim3(blocks), dim3(256), 0, st, p->C, p->S, p->U, p->V, p->uoff, p->soff,
                           p->rank, p->nt, p->len, p->glen, p->gsum, t1, t2, x, x1, x2, dst);
    }
    TRT_CHECK(hipGetLastError());
}

template <typename T>
T *upload(const T *h, size_t n, hipStream_t st, size_t *bytes)
{
    T *d = nullptr;
    TRT_CHECK(hipMalloc(reinterpret_cast<void **>(&d), std::max<size_t>(n, 1) * sizeof(T)));
    if (n) TRT_CHECK(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, st));
    if (bytes) *bytes += n * sizeof(T);
    return d;
}

void pod_free(trt_pod *p)
{
    for (void *q : {(void *)p->C, (void *)p->S, (void *)p->U, (void *)p->V, (void *)p->uoff, (void *)p->soff,
                    (void *)p->rank})
        if (q) (void)hipFree(q);
    delete p;
}

template <typename F>
int pguarded(trt_ctx *ctx, F &&f)
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

}  // namespace

namespace trt {

void bcg_split_if(Ctx &c, const std::string &name)
{
    if (name != "BCg_next") return;
    const int n = 2 * (c.nx + c.ny) * c.ng;
    hipLaunchKernelGGL(bcg_split_kernel, dim3((n + 255) / 256), dim3(256), 0, c.stream, c.arr("BCg_next"),
                       c.arr("Cg_L_next"), c.arr("Cg_B_next"), c.arr("Cg_R_next"), c.arr("Cg_T_next"), c.nx, c.ny, c.ng);
    TRT_CHECK(hipGetLastError());
}

}  // namespace trt

namespace {

void dmd_free(trt_dmd *d)
{
    for (void *m : {(void *)d->C, (void *)d->W, (void *)d->B_dev, (void *)d->E, (void *)d->woff, (void *)d->soff,
                    (void *)d->rank})
        if (m) (void)hipFree(m);
    delete d;
}

// a basis may only write a registered FP64 context array (not I_crn / I_crn_old,
// which are stored in the sweep's own layout and precision)
size_t target_size(trt_ctx *ctx, const char *name)
{
    const trt::Ctx &c = ctx->c;
    const size_t len = trt_array_size(ctx, name);
    if (len == 0 || trt::derived_size(c, name) || c.arrays.find(name) == c.arrays.end())
        throw trt::Error(std::string("no array ") + name + " to reconstruct into");
    return len;
}

template <typename T>
void unregister(std::vector<T *> &v, T *x)
{
    for (size_t k = 0; k < v.size(); ++k)
        if (v[k] == x) {
            v.erase(v.begin() + k);
            return;
        }
}

}  // namespace

namespace trt {

void rom_release_all(Ctx &c)
{
    for (trt_pod *p : c.pods) pod_free(p);
    for (trt_dmd *d : c.dmds) dmd_free(d);
    c.pods.clear();
    c.dmds.clear();
}

}  // namespace trt

extern "C" {

int trt_pod_create(trt_ctx *ctx, trt_pod **out, const char *name, int gsum, int nt, const int *rank,
                   const double *C, const double *S, const double *U, const double *V)
{
    return pguarded(ctx, [&](trt::Ctx &c) {
        if (!out || !name || !rank || !C || !S || !U || !V) throw trt::Error("null argument");
        *out = nullptr;
        const size_t len = target_size(ctx, name);
        if (gsum != 0 && gsum != 1) throw trt::Error("PODgsum must be 0 or 1");
        if (nt < 1) throw trt::Error("the basis needs at least one data time");
        const int ng = gsum ? 1 : c.ng;
        if (len % (size_t)ng) throw trt::Error(std::string(name) + " does not split into N_g equal group blocks");
        std::vector<int> rk(rank, rank + ng);
        for (int r : rk)
            if (r < 0) throw trt::Error("negative rank");
        const size_t glen = len / ng;
        std::vector<long long> uoff(ng);
        std::vector<int> soff(ng);
        size_t uo = 0, so = 0;
        for (int g = 0; g < ng; ++g) {
            uoff[g] = (long long)uo;
            soff[g] = (int)so;
            uo += glen * (size_t)rk[g];
            so += (size_t)rk[g];
        }
        trt_pod *p = new trt_pod;
        try {
            p->ctx = ctx; p->name = name; p->gsum = gsum; p->ng = ng; p->nt = nt; p->len = len; p->glen = glen;
            size_t b = 0;
            p->C = upload(C, len, c.stream, &b);
            p->S = upload(S, so, c.stream, &b);
            p->U = upload(U, uo, c.stream, &b);
            p->V = upload(V, so * (size_t)nt, c.stream, &b);
            p->uoff = upload(uoff.data(), uoff.size(), c.stream, &b);
            p->soff = upload(soff.data(), soff.size(), c.stream, &b);
            p->rank = upload(rk.data(), rk.size(), c.stream, &b);
            p->ubytes = uo * sizeof(double);
            TRT_CHECK(hipStreamSynchronize(c.stream));
        } catch (...) {
            pod_free(p);
            throw;
        }
        c.pods.push_back(p);
        *out = p;
    });
}

int trt_pod_reconstruct(trt_pod *p, int t1, int t2, double x, double x1, double x2)
{
    if (!p) return -1;
    return pguarded(p->ctx, [&](trt::Ctx &c) {
        if (t1 < 1 || t1 > p->nt || t2 < 1 || t2 > p->nt) throw trt::Error("data time index out of range");
        if (t1 != t2 && x2 == x1) throw trt::Error("interpolation between two data times at the same time");
        pod_launch(p, c.arr(p->name.c_str()), c.stream, t1 - 1, t2 - 1, x, x1, x2);
        trt::bcg_split_if(c, p->name);
        TRT_CHECK(hipStreamSynchronize(c.stream));
    });
}

int trt_pod_timed(trt_pod *p, int reps, double *ms)
{
    if (!p) return -1;
    return pguarded(p->ctx, [&](trt::Ctx &c) {
        if (reps < 1 || !ms) throw trt::Error("bad argument");
        double *dst = c.arr(p->name.c_str());
        TRT_CHECK(hipEventRecord(c.ev0, c.stream));
        for (int k = 0; k < reps; ++k) pod_launch(p, dst, c.stream, 0, 0, 0.0, 0.0, 0.0);
        TRT_CHECK(hipEventRecord(c.ev1, c.stream));
        TRT_CHECK(hipEventSynchronize(c.ev1));
        float f = 0.f;
        TRT_CHECK(hipEventElapsedTime(&f, c.ev0, c.ev1));
        *ms = f / reps;
    });
}

size_t trt_pod_basis_bytes(const trt_pod *p) { return p ? p->ubytes : 0; }

int trt_dmd_create(trt_ctx *ctx, trt_dmd **out, const char *name, int gsum, const int *rank, const double *C,
                   const double *L, const double *B, const double *W)
{
    return pguarded(ctx, [&](trt::Ctx &c) {
        if (!out || !name || !rank || !C || !L || !B || !W) throw trt::Error("null argument");
        *out = nullptr;
        const size_t len = target_size(ctx, name);
        if (gsum != 0 && gsum != 1) throw trt::Error("DMDgsum must be 0 or 1");
        const int ng = gsum ? 1 : c.ng;
        if (len % (size_t)ng) throw trt::Error(std::string(name) + " does not split into N_g equal group blocks");
        std::vector<int> rk(rank, rank + ng);
        for (int r : rk)
            if (r < 0) throw trt::Error("negative rank");
        const size_t glen = len / ng;
        std::vector<long long> woff(ng);
        std::vector<int> soff(ng);
        size_t wo = 0, so = 0;
        for (int g = 0; g < ng; ++g) {
            woff[g] = (long long)wo;
            soff[g] = (int)so;
            wo += glen * (size_t)rk[g];
            so += (size_t)rk[g];
        }
        trt_dmd *d = new trt_dmd;
        try {
            d->ctx = ctx; d->name = name; d->gsum = gsum; d->ng = ng; d->len = len; d->glen = glen;
            d->L.assign(L, L + 2 * so);
            d->B.assign(B, B + 2 * so);
            d->rank_h = rk;
            d->C = upload(C, len, c.stream, nullptr);
            d->W = upload(W, 2 * wo, c.stream, nullptr);
            d->B_dev = upload(B, 2 * so, c.stream, nullptr);
            d->E = upload(B, 2 * so, c.stream, nullptr);      // overwritten per reconstruction
            d->woff = upload(woff.data(), woff.size(), c.stream, nullptr);
            d->soff = upload(soff.data(), soff.size(), c.stream, nullptr);
            d->rank = upload(rk.data(), rk.size(), c.stream, nullptr);
            TRT_CHECK(hipStreamSynchronize(c.stream));
        } catch (...) {
            dmd_free(d);
            throw;
        }
        c.dmds.push_back(d);
        *out = d;
    });
}

int trt_dmd_reconstruct(trt_dmd *d, double time)
{
    if (!d) return -1;
    return pguarded(d->ctx, [&](trt::Ctx &c) {
        // exp(L t) per mode: exp(Re L t) (cos(Im L t) + i sin(Im L t))
        const size_t n = d->L.size() / 2;
        std::vector<double> e(2 * n);
        for (size_t k = 0; k < n; ++k) {
            const double m = std::exp(d->L[2 * k] * time), a = d->L[2 * k + 1] * time;
            e[2 * k] = m * std::cos(a);
            e[2 * k + 1] = m * std::sin(a);
        }
        if (n) TRT_CHECK(hipMemcpyAsync(d->E, e.data(), e.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
        const int blocks = (int)std::min<size_t>((d->len + 255) / 256, 8192);
        hipLaunchKernelGGL(dmd_kernel, dim3(blocks), dim3(256), 0, c.stream, d->C, reinterpret_cast<const double2 *>(d->W),
                           reinterpret_cast<const double2 *>(d->B_dev), reinterpret_cast<const double2 *>(d->E),
                           d->woff, d->soff, d->rank, d->len, d->glen, d->gsum, c.arr(d->name.c_str()));
        TRT_CHECK(hipGetLastError());
        trt::bcg_split_if(c, d->name);
        TRT_CHECK(hipStreamSynchronize(c.stream));
    });
}

int trt_dmd_destroy(trt_dmd *d)
{
    if (!d) return -1;
    (void)hipStreamSynchronize(d->ctx->c.stream);
    unregister(d->ctx->c.dmds, d);
    dmd_free(d);
    return 0;
}

int trt_pod_destroy(trt_pod *p)
{
    if (!p) return -1;
    (void)hipStreamSynchronize(p->ctx->c.stream);
    unregister(p->ctx->c.pods, p);
    pod_free(p);
    return 0;
}

}  // extern "C"
