// CONVERGENCE (MainSrc/CONVERGENCE_CHECKS.f08:7-44) on caller-owned device arrays
// without any allocation, for callers that cannot call hipMalloc: code under HIP
// graph capture, tight serving loops, harnesses that forbid it.
//
// The reduction's block partials (what trt_convergence keeps in a context's
// scratch) live in a file-scope device array, `scratch`, of NA_WORDS doubles
// (4 KiB + 32 B); every GPU the library runs on holds its own copy until the
// process ends.  Host code never passes the array's name to a kernel: it takes
// each device's address of its copy once (hipGetSymbolAddress, with that device
// current) and passes that address.  The result {norm, rho, converged} is written
// to the caller's device memory, so the call issues only two kernel launches on
// the caller's stream and never synchronises.  Every call on one GPU uses the same
// array: calls that overlap on one GPU (two streams, two host threads) race.
// Nothing in the array needs to start at zero: every partial a call reads was
// written by the same call.
#include "trt_internal.h"

#include <climits>
#include <mutex>

namespace {

constexpr int NA_BLOCKS = 128;                 // at most this many partial blocks
constexpr int NA_WORDS = 4 * NA_BLOCKS + 8;    // 4 partials per block, then the 4 combined maxima/minimum

static __device__ __align__(256) double scratch[NA_WORDS];

std::mutex g_mu;
std::vector<double *> g_scratch;               // per device: the address of that device's scratch

// the current device's copy of scratch (looked up at the first call on each device)
double *scratch_ptr()
{
    int dev = 0;
    TRT_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_mu);
    if ((size_t)dev >= g_scratch.size()) g_scratch.resize((size_t)dev + 1, nullptr);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw trt::Error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = static_cast<double *>(p);
    }
    return g_scratch[dev];
}

// Combines the block partials in block order (max and min are exact, so the result
// is that of trt_convergence's host-side combination) and forms the norm, rho and
// the test of CONVERGENCE_CHECKS.f08:20-43 with the same operations.
__global__ void conv_final(const double *part, int blocks, int type, double eps, int its, double *out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double a = 0.0, b = 0.0, cc = 0.0, mn = 1.7976931348623157e308;
    for (int k = 0; k < blocks; ++k) {
        a = fmax(a, part[k]);
        b = fmax(b, part[blocks + k]);
        cc = fmax(cc, part[2 * blocks + k]);
        mn = fmin(mn, part[3 * blocks + k]);
    }
    const double r = its > 1 ? a / b : 2.0;
    double nr = 0.0;
    if (type == 1) nr = a / cc;
    else if (type == 2) nr = cc;
    else nr = a / mn;
    const double new_eps = fabs(1.0 / r - 1.0) * eps;
    out[0] = nr;
    out[1] = r;
    out[2] = nr > new_eps ? 0.0 : 1.0;
}

}  // namespace

extern "C" int trt_convergence_noalloc(const double *x, const double *x_old, const double *x_old2, size_t n,
                                       int conv_type, double eps, int its, double *out, void *stream)
{
    try {
        if (!x || !x_old || !x_old2 || !out) throw trt::Error("null argument");
        if (n == 0) throw trt::Error("convergence arrays must not be empty");
        if (n > (size_t)INT_MAX)
            throw trt::Error("convergence arrays are limited to 2147483647 values (the reduction indexes them as int)");
        if (conv_type < 1 || conv_type > 3) throw trt::Error("conv_type must be 1, 2 or 3");
        double *sp = scratch_ptr();
        const hipStream_t s = static_cast<hipStream_t>(stream);
        const int blocks = trt::conv_blocks(n);
        if (blocks > NA_BLOCKS) throw trt::Error("reduction needs more partial blocks than the scratch array holds");
        trt::conv_partial_launch(x, x_old, x_old2, (int)n, conv_type, sp, blocks, s);
        hipLaunchKernelGGL(conv_final, dim3(1), dim3(64), 0, s, sp, blocks, conv_type, eps, its, out);
        TRT_CHECK(hipGetLastError());
        return 0;
    } catch (const trt::Error &e) {
        trt::set_thread_error(e.msg);
    } catch (const std::exception &e) {
        trt::set_thread_error(e.what());
    }
    return -2;
}
