// Autotuner (no counterpart in the reference): times the batched persistent
// sweep of a set of problems with every tile shape the tile engine offers for
// the precision -- and, for trt_autotune_batch, every column-segment length --
// and returns the fastest.
#include "trt_internal.h"

#include <limits>

namespace {

struct Shape {
    int tx, ty;
};

// tile shapes with a kernel in the tile engine, per precision (FP64 arithmetic:
// the FP64 shapes, also for FP32 storage)
const Shape kFp32[] = {{2, 2}, {2, 4}, {4, 2}, {4, 4}, {8, 2}, {2, 8}};
const Shape kFp64[] = {{2, 2}, {2, 4}, {4, 2}};

// the segment lengths trt_autotune_batch tries
const int kSegs[] = {1, 2, 4, 8};

// ms per batched sweep of the problems with one tile shape, for each of the
// nseg segment lengths segs[] (0 = the batch's own choice); contexts are
// created once per shape
void time_shape(const trt_problem *probs, int n, int precision, Shape sh, int sweeps, const int *segs, int nseg,
                double *ms)
{
    std::vector<trt_ctx *> ctxs;
    trt::PkBatch *b = nullptr;
    try {
        std::vector<trt::Ctx *> cs;
        for (int i = 0; i < n; ++i) {
            trt_problem p = probs[i];
            p.precision = precision;
            p.tile_x = sh.tx;
            p.tile_y = sh.ty;
            p.sweep_mode = 2;
            trt_ctx *c = nullptr;
            if (trt_ctx_create(&c, &p) != 0) throw trt::Error(std::string("autotune: ") + trt_last_error(nullptr));
            ctxs.push_back(c);
            cs.push_back(&c->c);
        }
        for (int s = 0; s < nseg; ++s) {
            b = trt::pk_batch_create(cs, false, segs[s]);
            trt::pk_batch_sweep(*b, 3);                 // warm-up
            trt::pk_batch_check(*b);
            for (int k = 0; k < sweeps; ++k) trt::pk_batch_sweep(*b, (k == 0 ? 1 : 0) | (k == sweeps - 1 ? 2 : 0));
            ms[s] = trt::pk_batch_elapsed_ms(*b) / sweeps;
            trt::pk_batch_check(*b);
            trt::pk_batch_destroy(b);
            b = nullptr;
        }
    } catch (...) {
        trt::pk_batch_destroy(b);
        for (trt_ctx *c : ctxs) trt_ctx_destroy(c);
        throw;
    }
    for (trt_ctx *c : ctxs) trt_ctx_destroy(c);
}

int shapes_of(int precision, const Shape **shapes)
{
    *shapes = precision == 1 ? kFp32 : kFp64;
    return precision == 1 ? (int)(sizeof(kFp32) / sizeof(Shape)) : (int)(sizeof(kFp64) / sizeof(Shape));
}

thread_local std::string g_tune_error;

}  // namespace

extern "C" {

int trt_autotune_tile(const trt_problem *probs, int n, int precision, int sweeps, int *tile_x, int *tile_y,
                      double *best_ms, double *all_ms)
{
    if (!probs || n < 1 || !tile_x || !tile_y || precision < 0 || precision > 2) return -1;
    try {
        const Shape *shapes = nullptr;
        const int ns = shapes_of(precision, &shapes);
        double best = std::numeric_limits<double>::infinity();
        const int automatic = 0;
        for (int s = 0; s < ns; ++s) {
            double ms = 0.0;
            time_shape(probs, n, precision, shapes[s], sweeps > 0 ? sweeps : 3, &automatic, 1, &ms);
            if (all_ms) all_ms[s] = ms;
            if (ms < best) {
                best = ms;
                *tile_x = shapes[s].tx;
                *tile_y = shapes[s].ty;
            }
        }
        if (best_ms) *best_ms = best;
        return 0;
    } catch (const trt::Error &e) {
        g_tune_error = e.msg;
    } catch (const std::exception &e) {
        g_tune_error = e.what();
    }
    return -2;
}

int trt_autotune_batch(const trt_problem *probs, int n, int precision, int sweeps, int *tile_x, int *tile_y,
                       int *seg, double *best_ms, double *all_ms)
{
    if (!probs || n < 1 || !tile_x || !tile_y || !seg || precision < 0 || precision > 2) return -1;
    try {
        const Shape *shapes = nullptr;
        const int ns = shapes_of(precision, &shapes);
        const int nseg = (int)(sizeof(kSegs) / sizeof(int));
        double best = std::numeric_limits<double>::infinity();
        for (int s = 0; s < ns; ++s) {
            double ms[sizeof(kSegs) / sizeof(int)];
            time_shape(probs, n, precision, shapes[s], sweeps > 0 ? sweeps : 3, kSegs, nseg, ms);
            for (int q = 0; q < nseg; ++q) {
                if (all_ms) all_ms[s * nseg + q] = ms[q];
                if (ms[q] < best) {
                    best = ms[q];
                    *tile_x = shapes[s].tx;
                    *tile_y = shapes[s].ty;
                    *seg = kSegs[q];
                }
            }
        }
        if (best_ms) *best_ms = best;
        return 0;
    } catch (const trt::Error &e) {
        g_tune_error = e.msg;
    } catch (const std::exception &e) {
        g_tune_error = e.what();
    }
    return -2;
}

int trt_autotune_segments(int *seg, int cap)
{
    const int nseg = (int)(sizeof(kSegs) / sizeof(int));
    for (int q = 0; q < nseg && q < cap; ++q)
        if (seg) seg[q] = kSegs[q];
    return nseg;
}

int trt_autotune_shapes(int precision, int *tile_x, int *tile_y, int cap)
{
    const Shape *shapes = precision == 1 ? kFp32 : kFp64;
    const int ns = precision == 1 ? (int)(sizeof(kFp32) / sizeof(Shape)) : (int)(sizeof(kFp64) / sizeof(Shape));
    for (int s = 0; s < ns && s < cap; ++s) {
        if (tile_x) tile_x[s] = shapes[s].tx;
        if (tile_y) tile_y[s] = shapes[s].ty;
    }
    return ns;
}

const char *trt_autotune_last_error(void) { return g_tune_error.c_str(); }

}  // extern "C"
