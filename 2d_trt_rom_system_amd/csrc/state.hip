// Device-resident state of one problem: allocation, the reference-named array
// registry, initialisation (MainSrc/INITIALIZERS.f08, ALGORITHMS.f08:203-218)
// and small device utilities.
#include "physics_device.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>

namespace trt {

void *Ctx::dalloc(size_t bytes)
{
    void *p = nullptr;
    size_t b = bytes ? bytes : 8;
    if (arena) {
        p = arena_alloc(arena, b);
        if (!p) throw Error("device arena exhausted");
    } else if (pooled) {
        p = cache_alloc(b);
    } else {
        TRT_CHECK(hipMalloc(&p, b));
    }
    owned.push_back(p);
    owned_bytes.push_back(b);
    this->bytes += b;
    return p;
}

void Ctx::dfree_last()
{
    if (owned.empty()) return;
    void *p = owned.back();
    const size_t b = owned_bytes.back();
    owned.pop_back();
    owned_bytes.pop_back();
    this->bytes -= b;
    if (arena) arena_release(arena, p, b);
    else if (pooled) cache_release(p, b);
    else (void)hipFree(p);
}

// ---------------------------------------------------------------------------
// device arena
// ---------------------------------------------------------------------------
struct Arena {
    char *base = nullptr;
    size_t size = 0, used = 0;
    std::map<size_t, size_t> free;      // offset -> length, address order, no two adjacent
    std::mutex m;
};

Arena *arena_create(size_t bytes)
{
    Arena *a = new Arena();
    a->size = bytes / 256 * 256;
    const hipError_t e = hipMalloc(reinterpret_cast<void **>(&a->base), a->size);
    if (e != hipSuccess) {
        delete a;
        throw Error(std::string("device arena of ") + std::to_string(bytes) + " bytes: " + hipGetErrorString(e));
    }
    a->free.emplace(0, a->size);
    return a;
}

void arena_destroy(Arena *a)
{
    if (!a) return;
    (void)hipFree(a->base);
    delete a;
}

void *arena_alloc(Arena *a, size_t &bytes)
{
    const size_t want = (std::max<size_t>(bytes, 1) + 255) / 256 * 256;
    std::lock_guard<std::mutex> lk(a->m);
    auto best = a->free.end();
    for (auto it = a->free.begin(); it != a->free.end(); ++it)
        if (it->second >= want && (best == a->free.end() || it->second < best->second)) best = it;
    if (best == a->free.end()) return nullptr;
    const size_t off = best->first, len = best->second;
    a->free.erase(best);
    // large blocks (intensity buffers) are cut from the top of the range, small
    // ones from the bottom, so the two kinds do not interleave and freed
    // contexts leave large contiguous ranges behind
    const bool large = want >= ((size_t)256 << 20);
    const size_t at = large ? off + len - want : off;
    if (len > want) a->free.emplace(large ? off : off + want, len - want);
    a->used += want;
    bytes = want;
    return a->base + at;
}

void arena_release(Arena *a, void *p, size_t bytes)
{
    std::lock_guard<std::mutex> lk(a->m);
    size_t off = static_cast<size_t>(static_cast<char *>(p) - a->base), len = bytes;
    a->used -= bytes;
    auto nx = a->free.lower_bound(off);
    if (nx != a->free.end() && off + len == nx->first) {       // merge with the following range
        len += nx->second;
        nx = a->free.erase(nx);
    }
    if (nx != a->free.begin()) {                              // merge with the preceding range
        auto pv = std::prev(nx);
        if (pv->first + pv->second == off) {
            pv->second += len;
            return;
        }
    }
    a->free.emplace(off, len);
}

size_t arena_used(const Arena *a) { return a ? a->used : 0; }
size_t arena_size(const Arena *a) { return a ? a->size : 0; }

// ---------------------------------------------------------------------------
// device block cache
// ---------------------------------------------------------------------------
namespace {
struct BlockCache {
    std::mutex m;
    std::multimap<size_t, void *> free;   // block size -> block
    size_t cached = 0;
    std::multimap<size_t, double *> pinned;   // pinned host buffers by capacity (doubles)
    std::vector<hipStream_t> streams;         // idle streams
};
BlockCache &block_cache()
{
    static BlockCache *bc = new BlockCache();   // outlives every context
    return *bc;
}
// blocks of 1 MiB and more are rounded to 2 MiB, smaller ones to 4 KiB
size_t round_block(size_t b)
{
    const size_t g = b >= ((size_t)1 << 20) ? ((size_t)2 << 20) : 4096;
    return (b + g - 1) / g * g;
}
}  // namespace

void *cache_alloc(size_t &bytes)
{
    const size_t want = round_block(bytes);
    BlockCache &bc = block_cache();
    {
        std::lock_guard<std::mutex> lk(bc.m);
        auto it = bc.free.lower_bound(want);
        // reuse a cached block at most 1/16 larger than asked for
        if (it != bc.free.end() && it->first <= want + want / 16) {
            void *p = it->second;
            bytes = it->first;
            bc.cached -= it->first;
            bc.free.erase(it);
            return p;
        }
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, want);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        cache_trim();
        e = hipMalloc(&p, want);
    }
    if (e != hipSuccess) throw Error(std::string("hipMalloc: ") + hipGetErrorString(e));
    bytes = want;
    return p;
}

void cache_release(void *p, size_t bytes)
{
    BlockCache &bc = block_cache();
    std::lock_guard<std::mutex> lk(bc.m);
    bc.free.emplace(bytes, p);
    bc.cached += bytes;
}

void cache_trim()
{
    BlockCache &bc = block_cache();
    std::lock_guard<std::mutex> lk(bc.m);
    for (auto &kv : bc.free) (void)hipFree(kv.second);
    bc.free.clear();
    bc.cached = 0;
    for (auto &kv : bc.pinned) (void)hipHostFree(kv.second);
    bc.pinned.clear();
    for (hipStream_t s : bc.streams) (void)hipStreamDestroy(s);
    bc.streams.clear();
}

double *cache_pinned_get(size_t n)
{
    BlockCache &bc = block_cache();
    {
        std::lock_guard<std::mutex> lk(bc.m);
        auto it = bc.pinned.lower_bound(n);
        if (it != bc.pinned.end() && it->first == n) {
            double *p = it->second;
            bc.pinned.erase(it);
            return p;
        }
    }
    double *p = nullptr;
    TRT_CHECK(hipHostMalloc(reinterpret_cast<void **>(&p), n * sizeof(double), hipHostMallocDefault));
    return p;
}

void cache_pinned_put(double *p, size_t n)
{
    BlockCache &bc = block_cache();
    std::lock_guard<std::mutex> lk(bc.m);
    bc.pinned.emplace(n, p);
}

hipStream_t cache_stream_get()
{
    BlockCache &bc = block_cache();
    {
        std::lock_guard<std::mutex> lk(bc.m);
        if (!bc.streams.empty()) {
            hipStream_t s = bc.streams.back();
            bc.streams.pop_back();
            return s;
        }
    }
    hipStream_t s = nullptr;
    TRT_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    return s;
}

void cache_stream_put(hipStream_t s)
{
    BlockCache &bc = block_cache();
    std::lock_guard<std::mutex> lk(bc.m);
    bc.streams.push_back(s);
}

size_t cache_bytes()
{
    BlockCache &bc = block_cache();
    std::lock_guard<std::mutex> lk(bc.m);
    return bc.cached;
}

size_t state_bytes(const trt_problem &p)
{
    const size_t na = (size_t)p.nx * p.ny, nv = (size_t)(p.nx + 1) * p.ny, nh = (size_t)p.nx * (p.ny + 1);
    const size_t nbd = (size_t)p.nx + p.ny;
    size_t n = (27 * na + 14 * nv + 14 * nh + 10 * nbd) * p.ng;      // per-group arrays of init_state
    n += 40 * na + 6 * nv + 6 * nh + 6 * nbd + 3 * (size_t)p.nm + p.ng + 1 + p.nmat + p.nx + p.ny;
    n += std::max<size_t>(4096, 8 * (size_t)std::max(p.nx, p.ny) * p.ng + 4096) + 4 * (size_t)p.ng;   // scratch, bcin
    return n * 8 + na * sizeof(int);
}

DevArray &Ctx::reg(const std::string &name, int d0, int d1, int d2, int d3)
{
    DevArray a;
    a.d[0] = d0; a.d[1] = d1; a.d[2] = d2; a.d[3] = d3;
    a.p = static_cast<double *>(dalloc(a.size() * sizeof(double)));
    TRT_CHECK(hipMemsetAsync(a.p, 0, a.size() * sizeof(double), stream));
    arrays[name] = a;
    return arrays[name];
}

double *Ctx::arr(const char *name)
{
    auto it = arrays.find(name);
    if (it == arrays.end()) throw Error(std::string("unknown array ") + name);
    return it->second.p;
}

// ---------------------------------------------------------------------------
// small utilities
// ---------------------------------------------------------------------------
__global__ void fill_kernel(double *x, double v, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) x[i] = v;
}

static int nb(size_t n) { size_t b = (n + 255) / 256; return (int)std::min<size_t>(std::max<size_t>(b, 1), 65536); }

void dfill(Ctx &c, double *dst, double v, size_t n)
{
    hipLaunchKernelGGL(fill_kernel, dim3(nb(n)), dim3(256), 0, c.stream, dst, v, n);
}

void dcopy(Ctx &c, double *dst, const double *src, size_t n)
{
    TRT_CHECK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
}

// max |x| over n values, single block-strided pass + final pass (deterministic)
__global__ void maxabs_partial(const double *x, size_t n, double *part)
{
    __shared__ double s[256];
    double m = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        m = fmax(m, fabs(x[i]));
    s[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

double dmaxabs(Ctx &c, const double *x, size_t n)
{
    const int blocks = std::min(nb(n), 256);
    hipLaunchKernelGGL(maxabs_partial, dim3(blocks), dim3(256), 0, c.stream, x, n, c.scratch);
    hipLaunchKernelGGL(maxabs_partial, dim3(1), dim3(256), 0, c.stream, c.scratch, (size_t)blocks, c.scratch + 256);
    TRT_CHECK(hipMemcpyAsync(c.hbuf, c.scratch + 256, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    TRT_CHECK(hipStreamSynchronize(c.stream));
    return c.hbuf[0];
}

// ---------------------------------------------------------------------------
// CONVERGENCE, CONVERGENCE_CHECKS.f08:7-44: the three max-reductions are done in
// one pass; type 3 needs min|x|.
// ---------------------------------------------------------------------------
__global__ void conv_partial(const double *x, const double *xo, const double *xo2, int n, int type, double *part)
{
    __shared__ double s[4][256];
    double a = 0.0, b = 0.0, cc = 0.0, mn = 1.7976931348623157e308;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        a = fmax(a, fabs(x[i] - xo[i]));
        b = fmax(b, fabs(xo[i] - xo2[i]));
        if (type == 2) cc = fmax(cc, fabs(1.0 - xo[i] / x[i]));
        else cc = fmax(cc, fabs(x[i]));
        mn = fmin(mn, fabs(x[i]));
    }
    s[0][threadIdx.x] = a; s[1][threadIdx.x] = b; s[2][threadIdx.x] = cc; s[3][threadIdx.x] = mn;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            for (int k = 0; k < 3; ++k) s[k][threadIdx.x] = fmax(s[k][threadIdx.x], s[k][threadIdx.x + o]);
            s[3][threadIdx.x] = fmin(s[3][threadIdx.x], s[3][threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) for (int k = 0; k < 4; ++k) part[k * gridDim.x + blockIdx.x] = s[k][0];
}

int conv_blocks(size_t n) { return std::min(nb(n), 128); }

void conv_partial_launch(const double *x, const double *xo, const double *xo2, int n, int conv_type, double *part,
                         int blocks, hipStream_t s)
{
    hipLaunchKernelGGL(conv_partial, dim3(blocks), dim3(256), 0, s, x, xo, xo2, n, conv_type, part);
}

void convergence(Ctx &c, const double *x, const double *xo, const double *xo2, int n, int conv_type, double eps,
                 int its, double *norm, double *rho, int *conv)
{
    const int blocks = conv_blocks((size_t)n);
    conv_partial_launch(x, xo, xo2, n, conv_type, c.scratch, blocks, c.stream);
    TRT_CHECK(hipMemcpyAsync(c.hbuf, c.scratch, sizeof(double) * 4 * blocks, hipMemcpyDeviceToHost, c.stream));
    TRT_CHECK(hipStreamSynchronize(c.stream));
    double a = 0, b = 0, cc = 0, mn = 1.7976931348623157e308;
    for (int k = 0; k < blocks; ++k) {
        a = std::max(a, c.hbuf[k]);
        b = std::max(b, c.hbuf[blocks + k]);
        cc = std::max(cc, c.hbuf[2 * blocks + k]);
        mn = std::min(mn, c.hbuf[3 * blocks + k]);
    }
    const double r = its > 1 ? a / b : 2.0;
    double nr = 0.0;
    if (conv_type == 1) nr = a / cc;
    else if (conv_type == 2) nr = cc;
    else if (conv_type == 3) nr = a / mn;
    if (norm) *norm = nr;
    if (rho) *rho = r;
    const double new_eps = std::fabs(1.0 / r - 1.0) * eps;
    if (conv) *conv = nr > new_eps ? 0 : 1;
}

// ---------------------------------------------------------------------------
// initialisation
// ---------------------------------------------------------------------------
// incoming boundary intensities bcin[side][g] = Bg_planck(bcT(side)), sides L,B,R,T
// (RT_BC_UPDATE with BC_Type 0, UPDATES.f08:26-102)
__global__ void bc_kernel(double *bcin, const double *nu, double b0, double b1, double b2, double b3, int ng, double cu)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 4 * ng) return;
    const int side = idx / ng, g = idx % ng;
    const double T = side == 0 ? b0 : side == 1 ? b1 : side == 2 ? b2 : b3;
    bcin[idx] = d_bg_planck(T, nu[g], nu[g + 1], cu);
}

// per-group equilibrium values bg(Tini) (IN:86,268,392)
__global__ void bg_init_kernel(double *bg, const double *nu, double T, int ng, double cu)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < ng) bg[g] = d_bg_planck(T, nu[g], nu[g + 1], cu);
}

// value[i + n*g] = scale * bg[g]
__global__ void fill_group_kernel(double *x, const double *bg, double scale, size_t n, int ng)
{
    const size_t tot = n * ng;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x)
        x[i] = scale * bg[i / n];
}

// grey value[i] = sum_g scale*bg[g]
__global__ void fill_greysum_kernel(double *x, const double *bg, double scale, size_t n, int ng)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        double t = 0.0;
        for (int g = 0; g < ng; ++g) t += scale * bg[g];
        x[i] = t;
    }
}

template <typename TI>
__global__ void icrn_init_kernel(TI *I, const double *bg, size_t ncell, int ng, int K)
{
    const size_t NL = (size_t)ng * K;
    const size_t tot = ncell * 4 * NL * 4;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
        const size_t L = (i / 4) % NL;
        I[i] = (TI)(bg[L / K] * iscale<TI>());
    }
}

void init_state(Ctx &c)
{
    const int nx = c.nx, ny = c.ny, ng = c.ng, nm = c.nm;
    const size_t na = (size_t)nx * ny, nv = (size_t)(nx + 1) * ny, nh = (size_t)nx * (ny + 1);
    // register the reference-named arrays
    const char *cellg[] = {"Hg_avg_xx", "Hg_avg_yy", "Hg_avg_xy", "HO_Eg_avg", "Eg_avg", "Eg_avg_old", "fg_avg_xx",
                           "fg_avg_yy", "fg_avg_xy", "fg_avg_xx_old", "fg_avg_yy_old", "G_old", "Pold_L", "Pold_B",
                           "Pold_R", "Pold_T", "RT_Src", "MGQD_Src", "MGQD_Src_old", "KapE", "KapB", "KapR",
                           "KapE_old", "KapR_old", "Bg"};
    for (auto n : cellg) c.reg(n, nx, ny, ng);
    const char *vg[] = {"Hg_edgV_xx", "Hg_edgV_xy", "HO_Eg_edgV", "HO_Fxg_edgV", "Eg_edgV", "Eg_edgV_old", "Fxg_edgV",
                        "Fxg_edgV_old", "fg_edgV_xx", "fg_edgV_xy", "fg_edgV_xx_old", "fg_edgV_xy_old"};
    for (auto n : vg) c.reg(n, nx + 1, ny, ng);
    const char *hg[] = {"Hg_edgH_yy", "Hg_edgH_xy", "HO_Eg_edgH", "HO_Fyg_edgH", "Eg_edgH", "Eg_edgH_old", "Fyg_edgH",
                        "Fyg_edgH_old", "fg_edgH_yy", "fg_edgH_xy", "fg_edgH_yy_old", "fg_edgH_xy_old"};
    for (auto n : hg) c.reg(n, nx, ny + 1, ng);
    for (auto n : {"Cg_L", "Cg_R", "Eg_in_L", "Eg_in_R", "Fg_in_L", "Fg_in_R"}) c.reg(n, ny, ng);
    for (auto n : {"Cg_B", "Cg_T", "Eg_in_B", "Eg_in_T", "Fg_in_B", "Fg_in_T"}) c.reg(n, nx, ng);
    // the next time level's QD factors and C_g of the ROM runs' step, pushed by the caller
    for (auto n : {"fg_avg_xx_next", "fg_avg_yy_next"}) c.reg(n, nx, ny, ng);
    for (auto n : {"fg_edgV_xx_next", "fg_edgV_xy_next"}) c.reg(n, nx + 1, ny, ng);
    for (auto n : {"fg_edgH_yy_next", "fg_edgH_xy_next"}) c.reg(n, nx, ny + 1, ng);
    for (auto n : {"Cg_L_next", "Cg_R_next"}) c.reg(n, ny, ng);
    for (auto n : {"Cg_B_next", "Cg_T_next"}) c.reg(n, nx, ng);
    // C_g of all four sides as POD/DMD_RECONSTRUCT_BCg hold it, per group [L | B | R | T]
    c.reg("BCg_next", 2 * (nx + ny), ng);
    const char *grey[] = {"HO_E_avg", "MGQD_E_avg", "Temp", "Temp_Old", "E_avg", "E_avg_old", "KapE_Bar",
                          "KapE_Bar_old", "KapE_Bar_dT", "KapE_bar_MGQDold", "GQD_Src", "GQD_Src_old", "Gold_Hat",
                          "Rhat_old", "DC_xx", "DL_xx", "DR_xx", "DC_yy", "DB_yy", "DT_yy", "DL_xy", "DR_xy", "DB_xy",
                          "DT_xy", "PL", "PR", "PB", "PT", "Temp_RTold", "Temp_RTold2", "Temp_MGQDold",
                          "Temp_MGQDold2", "HO_E_avg_RTold", "HO_E_avg_RTold2", "MGQD_E_avg_MGQDold",
                          "MGQD_E_avg_MGQDold2", "A", "_conv_x", "_conv_x_old", "_conv_x_old2"};
    for (auto n : grey) c.reg(n, nx, ny);
    for (auto n : {"HO_E_edgV", "HO_Fx_edgV", "MGQD_E_edgV", "MGQD_Fx_edgV", "E_edgV", "Fx_edgV"}) c.reg(n, nx + 1, ny);
    for (auto n : {"HO_E_edgH", "HO_Fy_edgH", "MGQD_E_edgH", "MGQD_Fy_edgH", "E_edgH", "Fy_edgH"}) c.reg(n, nx, ny + 1);
    for (auto n : {"Cb_L", "Cb_R", "E_in_L", "E_in_R", "F_in_L", "F_in_R"}) c.reg(n, ny);
    for (auto n : {"Cb_B", "Cb_T", "E_in_B", "E_in_T", "F_in_B", "F_in_T"}) c.reg(n, nx);
    c.dx = c.reg("Delx", nx).p; c.dy = c.reg("Dely", ny).p;
    c.ox = c.reg("Omega_x", nm).p; c.oy = c.reg("Omega_y", nm).p; c.w = c.reg("quad_weight", nm).p;
    c.nu = c.reg("nu_g", ng + 1).p; c.kmult = c.reg("Kappa_Mult", c.nmat).p;
    c.A = c.arr("A");
    c.bcin = static_cast<double *>(c.dalloc(4 * (size_t)ng * sizeof(double)));
    c.mat = static_cast<int *>(c.dalloc(na * sizeof(int)));
    TRT_CHECK(hipMemcpyAsync(c.dx, c.p.dx, nx * 8, hipMemcpyHostToDevice, c.stream));
    TRT_CHECK(hipMemcpyAsync(c.dy, c.p.dy, ny * 8, hipMemcpyHostToDevice, c.stream));
    TRT_CHECK(hipMemcpyAsync(c.ox, c.p.ox, nm * 8, hipMemcpyHostToDevice, c.stream));
    TRT_CHECK(hipMemcpyAsync(c.oy, c.p.oy, nm * 8, hipMemcpyHostToDevice, c.stream));
    TRT_CHECK(hipMemcpyAsync(c.w, c.p.w, nm * 8, hipMemcpyHostToDevice, c.stream));
    TRT_CHECK(hipMemcpyAsync(c.nu, c.p.nu, (ng + 1) * 8, hipMemcpyHostToDevice, c.stream));
    TRT_CHECK(hipMemcpyAsync(c.kmult, c.p.kmult, c.nmat * 8, hipMemcpyHostToDevice, c.stream));
    TRT_CHECK(hipMemcpyAsync(c.mat, c.p.mat, na * sizeof(int), hipMemcpyHostToDevice, c.stream));
    {
        std::vector<double> A(na);
        for (int j = 0; j < ny; ++j) for (int i = 0; i < nx; ++i) A[i + (size_t)nx * j] = c.p.dx[i] * c.p.dy[j];  // MISC_INIT
        TRT_CHECK(hipMemcpyAsync(c.A, A.data(), na * 8, hipMemcpyHostToDevice, c.stream));
        TRT_CHECK(hipStreamSynchronize(c.stream));
    }
    c.scratch_n = std::max<size_t>(4096, 8 * (size_t)std::max(nx, ny) * ng + 4096);
    c.scratch = static_cast<double *>(c.dalloc(c.scratch_n * sizeof(double)));
    if (c.pooled) c.hbuf = cache_pinned_get(4096);
    else TRT_CHECK(hipHostMalloc(reinterpret_cast<void **>(&c.hbuf), 4096 * sizeof(double), hipHostMallocDefault));

    sweep_setup(c);

    const double cu = c.p.comp_unit;
    hipLaunchKernelGGL(bc_kernel, dim3((4 * ng + 255) / 256), dim3(256), 0, c.stream, c.bcin, c.nu, c.p.bcT[0],
                       c.p.bcT[1], c.p.bcT[2], c.p.bcT[3], ng, cu);
    double *bg = c.scratch + 1024;   // ng values
    hipLaunchKernelGGL(bg_init_kernel, dim3((ng + 255) / 256), dim3(256), 0, c.stream, bg, c.nu, c.p.tini, ng, cu);
    c.h_bcin.resize(4 * (size_t)ng);
    TRT_CHECK(hipMemcpyAsync(c.h_bcin.data(), c.bcin, 4 * ng * 8, hipMemcpyDeviceToHost, c.stream));
    // RT_INIT sets the inflow of every side, reflective ones included, to Bg_planck(bcT)
    // (RT_BC_UPDATE with BC_Type (0,0,0,0), INITIALIZERS.f08:106)
    reflect_reset(c);

    // RT_INIT, IN:81-107: all intensities at bg(Tini); boundary values are the bcin array
    const size_t nrec = na * 4 * (size_t)ng * c.sg.K * 4;
    for (int b = 0; b < 2; ++b) {
        if (c.precision)
            hipLaunchKernelGGL(icrn_init_kernel<float>, dim3(nb(nrec)), dim3(256), 0, c.stream,
                               static_cast<float *>(c.Icrn[b]), bg, na, ng, c.sg.K);
        else
            hipLaunchKernelGGL(icrn_init_kernel<double>, dim3(nb(nrec)), dim3(256), 0, c.stream,
                               static_cast<double *>(c.Icrn[b]), bg, na, ng, c.sg.K);
    }
    c.icur = 0; c.iold = 0;
    const double e = 4.0 * PI / c.p.c;
    auto gfill = [&](const char *name, size_t n) {
        hipLaunchKernelGGL(fill_group_kernel, dim3(nb(n * ng)), dim3(256), 0, c.stream, c.arr(name), bg, e, n, ng);
    };
    auto gsum = [&](const char *name, size_t n) {
        hipLaunchKernelGGL(fill_greysum_kernel, dim3(nb(n)), dim3(256), 0, c.stream, c.arr(name), bg, e, n, ng);
    };
    gfill("HO_Eg_avg", na); gfill("HO_Eg_edgV", nv); gfill("HO_Eg_edgH", nh);
    gsum("HO_E_avg", na); gsum("HO_E_edgV", nv); gsum("HO_E_edgH", nh);
    // MGQD_INIT, IN:263-319
    gfill("Eg_avg", na); gfill("Eg_edgV", nv); gfill("Eg_edgH", nh);
    gsum("MGQD_E_avg", na); gsum("MGQD_E_edgV", nv); gsum("MGQD_E_edgH", nh);
    dfill(c, c.arr("fg_avg_xx"), 1.0 / 3.0, na * ng);
    dfill(c, c.arr("fg_avg_yy"), 1.0 / 3.0, na * ng);
    dfill(c, c.arr("fg_edgV_xx"), 1.0 / 3.0, nv * ng);
    dfill(c, c.arr("fg_edgH_yy"), 1.0 / 3.0, nh * ng);
    dfill(c, c.arr("Cg_L"), c.p.bc_type[0] == 0 ? -0.5 : 0.0, (size_t)ny * ng);
    dfill(c, c.arr("Cg_B"), c.p.bc_type[1] == 0 ? -0.5 : 0.0, (size_t)nx * ng);
    dfill(c, c.arr("Cg_R"), c.p.bc_type[2] == 0 ? 0.5 : 0.0, (size_t)ny * ng);
    dfill(c, c.arr("Cg_T"), c.p.bc_type[3] == 0 ? 0.5 : 0.0, (size_t)nx * ng);
    mgqd_in_calc(c);
    // GQD_INIT, IN:387-425
    gsum("E_avg", na); gsum("E_edgV", nv); gsum("E_edgH", nh);
    dfill(c, c.arr("Cb_L"), c.p.bc_type[0] == 0 ? -0.5 : 0.0, ny);
    dfill(c, c.arr("Cb_B"), c.p.bc_type[1] == 0 ? -0.5 : 0.0, nx);
    dfill(c, c.arr("Cb_R"), c.p.bc_type[2] == 0 ? 0.5 : 0.0, ny);
    dfill(c, c.arr("Cb_T"), c.p.bc_type[3] == 0 ? 0.5 : 0.0, nx);
    gqd_in_calc(c);
    // TEMP_INIT, IN:136-137
    dfill(c, c.arr("Temp"), c.p.tini, na);
    material_update(c);
    TRT_CHECK(hipStreamSynchronize(c.stream));
    c.time = 0.0;
}

}  // namespace trt
