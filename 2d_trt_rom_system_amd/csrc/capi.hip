// The extern "C" boundary (include/trt_capi.h): argument checking, error
// capture and the host side of every entry point.
#include "trt_internal.h"

#include <cstdlib>
#include <cstring>

using trt::Ctx;
using trt::Error;

struct trt_batch {
    trt::PkBatch *b = nullptr;
    std::string err;
};

namespace {

template <typename F>
int guarded(trt_ctx *ctx, F &&f)
{
    if (!ctx) return -1;
    try {
        f(ctx->c);
        return 0;
    } catch (const Error &e) {
        ctx->c.err = e.msg;
    } catch (const std::exception &e) {
        ctx->c.err = e.what();
    }
    return -2;
}

thread_local std::string g_create_error;

void validate(const trt_problem *p)
{
    if (!p) throw Error("null problem");
    if (p->nx < 2 || p->ny < 2) throw Error("N_x and N_y must be at least 2 (QD_SOLVERS.f08 couples neighbouring cells)");
    if (p->nm < 4 || p->nm % 4) throw Error("N_m must be a positive multiple of 4");
    if (p->ng < 1) throw Error("N_g must be positive");
    if (p->nmat < 1) throw Error("need at least one material");
    if (!p->dx || !p->dy || !p->ox || !p->oy || !p->w || !p->nu || !p->mat || !p->kmult)
        throw Error("null problem vector");
    // BC_Type 2 passes the reference's input check (INPUTS.f08:1024-1027) but no routine
    // gives it a boundary condition: RT_BC_UPDATE, Cg_Calc and MGQD_INIT branch on 0 and 1
    for (int k = 0; k < 4; ++k)
        if (p->bc_type[k] != 0 && p->bc_type[k] != 1)
            throw Error("BC_Type must be 0 (incoming Planckian) or 1 (reflective); the reference defines no other");
    if (p->sweep_mode == 1)
        for (int k = 0; k < 4; ++k)
            if (p->bc_type[k] == 1)
                throw Error("reflective boundaries (BC_Type 1) run on the tile engine (sweep_mode 0, 2, 3 or 4)");
    if (p->precision < 0 || p->precision > 2)
        throw Error("precision must be 0 (FP64), 1 (FP32 sweep) or 2 (FP32 storage, FP64 arithmetic)");
    if (p->precision == 2 && p->sweep_mode == 1)
        throw Error("precision 2 runs on the tile engine (sweep_mode 0, 2, 3 or 4)");
    if (p->sweep_mode < 0 || p->sweep_mode > 4) throw Error("sweep_mode must be 0, 1, 2, 3 or 4");
    if (p->grey_emission != 0 && p->grey_emission != 1) throw Error("grey_emission must be 0 or 1");
    if (p->no_grey != 0 && p->no_grey != 1) throw Error("no_grey must be 0 or 1");
    if (p->use_safety_search && !(p->line_src > 1.0))
        throw Error("the safety search needs a line-search factor line_src above 1");
    if (p->qd_solver < 0 || p->qd_solver > 3)
        throw Error("qd_solver must be 0 (nested dissection, factors reused within a step), 1 (BiCGSTAB), 2 (block "
                    "LU) or 3 (nested dissection factorizing every solve)");
    if (p->dt <= 0 || p->c <= 0 || p->theta <= 0) throw Error("c, dt and theta must be positive");
    for (size_t k = 0; k < (size_t)p->nx * p->ny; ++k)
        if (p->mat[k] < 0 || p->mat[k] >= p->nmat) throw Error("material id out of range");
}

}  // namespace

extern "C" {

}  // extern "C"

namespace trt {

void set_thread_error(const std::string &msg) { g_create_error = msg; }

trt_ctx *ctx_create(const trt_problem *p, bool pooled, Arena *arena)
{
    trt_ctx *ctx = nullptr;
    try {
        validate(p);
        ctx = new trt_ctx();
        Ctx &c = ctx->c;
        c.pooled = pooled;
        c.arena = arena;
        c.p = *p;
        c.nx = p->nx; c.ny = p->ny; c.nm = p->nm; c.ng = p->ng; c.nmat = p->nmat;
        c.precision = p->precision;
        c.lin_rtol = p->lin_rtol > 0 ? p->lin_rtol : 1e-13;
        c.lin_maxit = p->lin_maxit > 0 ? p->lin_maxit : 2000;
        c.qd_solver = p->qd_solver;
        if (const char *v = getenv("TRT_VERBOSE")) c.verbose = atoi(v);
        if (pooled) c.stream = cache_stream_get();
        else TRT_CHECK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
        TRT_CHECK(hipEventCreate(&c.ev0));
        TRT_CHECK(hipEventCreate(&c.ev1));
        trt::init_state(c);
        // the problem's host vectors are not kept
        c.p.dx = c.p.dy = c.p.ox = c.p.oy = c.p.w = c.p.nu = c.p.kmult = nullptr;
        c.p.mat = nullptr;
        return ctx;
    } catch (...) {
        if (ctx) ctx_destroy(ctx);
        throw;
    }
}

void ctx_destroy(trt_ctx *ctx)
{
    Ctx &c = ctx->c;
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    trt::rom_release_all(c);
    trt::romgrid_release_all(c);
    trt::pk_destroy(c);
    trt::direct_destroy(c);
    trt::nd_destroy(c);
    for (size_t k = 0; k < c.owned.size(); ++k) {
        if (c.arena) arena_release(c.arena, c.owned[k], c.owned_bytes[k]);
        else if (c.pooled) cache_release(c.owned[k], c.owned_bytes[k]);
        else (void)hipFree(c.owned[k]);
    }
    if (c.hbuf) {
        if (c.pooled) cache_pinned_put(c.hbuf, 4096);
        else (void)hipHostFree(c.hbuf);
    }
    if (c.ev0) (void)hipEventDestroy(c.ev0);
    if (c.ev1) (void)hipEventDestroy(c.ev1);
    if (c.stream) {
        if (c.pooled) cache_stream_put(c.stream);
        else (void)hipStreamDestroy(c.stream);
    }
    delete ctx;
}

}  // namespace trt

extern "C" {

int trt_ctx_create(trt_ctx **out, const trt_problem *p)
{
    if (!out) return -1;
    *out = nullptr;
    try {
        *out = trt::ctx_create(p, false);
        return 0;
    } catch (const Error &e) {
        g_create_error = e.msg;
    } catch (const std::exception &e) {
        g_create_error = e.what();
    }
    return -2;
}

int trt_ctx_destroy(trt_ctx *ctx)
{
    if (!ctx) return -1;
    trt::ctx_destroy(ctx);
    return 0;
}

const char *trt_last_error(const trt_ctx *ctx)
{
    if (!ctx) return g_create_error.c_str();
    return ctx->c.err.c_str();
}

size_t trt_ctx_bytes(const trt_problem *p)
{
    if (!p) return 0;
    return trt::sweep_bytes(*p, nullptr) + trt::state_bytes(*p);
}

size_t trt_ctx_device_bytes(const trt_ctx *ctx) { return ctx ? ctx->c.bytes : 0; }

size_t trt_array_size(const trt_ctx *ctx, const char *name)
{
    if (!ctx || !name) return 0;
    const Ctx &c = ctx->c;
    if (!strcmp(name, "I_crn") || !strcmp(name, "I_crn_old")) return (size_t)4 * c.nx * c.ny * c.nm * c.ng;
    if (const size_t d = trt::derived_size(c, name)) return d;
    auto it = c.arrays.find(name);
    return it == c.arrays.end() ? 0 : it->second.size();
}

int trt_push(trt_ctx *ctx, const char *name, const double *host, size_t n)
{
    return guarded(ctx, [&](Ctx &c) {
        if (!name || !host) throw Error("null argument");
        if (n != trt_array_size(ctx, name)) throw Error(std::string("size mismatch for ") + name);
        if (!strcmp(name, "I_crn")) { trt::push_icrn(c, c.icur, host); return; }
        if (!strcmp(name, "I_crn_old")) {
            // a pushed I_crn_old gets its own buffer when it shares one with I_crn
            if (c.iold == c.icur) c.iold = 1 - c.icur;
            trt::push_icrn(c, c.iold, host);
            return;
        }
        if (trt::derived_size(c, name)) throw Error(std::string(name) + " is derived from I_crn and cannot be pushed");
        TRT_CHECK(hipMemcpyAsync(c.arr(name), host, n * 8, hipMemcpyHostToDevice, c.stream));
        // the sweep reads KapE and RT_Src through its per-(cell, group) pair array
        if (!strcmp(name, "KapE") || !strcmp(name, "RT_Src")) trt::rebuild_sc(c);
        // a C_g boundary vector reconstructed by the caller reaches the step through Cg_*_next
        trt::bcg_split_if(c, name);
        TRT_CHECK(hipStreamSynchronize(c.stream));
    });
}

int trt_pull(trt_ctx *ctx, const char *name, double *host, size_t n)
{
    return guarded(ctx, [&](Ctx &c) {
        if (!name || !host) throw Error("null argument");
        if (n != trt_array_size(ctx, name)) throw Error(std::string("size mismatch for ") + name);
        if (!strcmp(name, "I_crn")) { trt::pull_icrn(c, c.icur, host); return; }
        if (!strcmp(name, "I_crn_old")) { trt::pull_icrn(c, c.iold, host); return; }
        if (trt::derived_size(c, name)) { trt::pull_derived(c, name, host); return; }
        TRT_CHECK(hipMemcpyAsync(host, c.arr(name), n * 8, hipMemcpyDeviceToHost, c.stream));
        TRT_CHECK(hipStreamSynchronize(c.stream));
    });
}

#define SIMPLE(fn, call)                                                          \
    int fn(trt_ctx *ctx)                                                          \
    {                                                                             \
        return guarded(ctx, [&](Ctx &c) { call(c); TRT_CHECK(hipStreamSynchronize(c.stream)); }); \
    }

SIMPLE(trt_material_update, trt::material_update)
SIMPLE(trt_collapse_intensities, trt::collapse_intensities)
SIMPLE(trt_fg_calc, trt::fg_calc)
SIMPLE(trt_cg_calc, trt::cg_calc)
SIMPLE(trt_mgqd_in_calc, trt::mgqd_in_calc)
SIMPLE(trt_gqd_in_calc, trt::gqd_in_calc)
SIMPLE(trt_old_mgqd_coefs, trt::old_mgqd_coefs)
SIMPLE(trt_old_grey_coefs, trt::old_grey_coefs)
SIMPLE(trt_collapse_mg_ef, trt::collapse_mg_ef)
SIMPLE(trt_cbar_calc, trt::cbar_calc)
SIMPLE(trt_grey_coefs, trt::grey_coefs)

int trt_transport_scb(trt_ctx *ctx)
{
    return guarded(ctx, [&](Ctx &c) {
        TRT_CHECK(hipEventRecord(c.ev0, c.stream));
        trt::transport_sweep(c);
        TRT_CHECK(hipEventRecord(c.ev1, c.stream));
        TRT_CHECK(hipEventSynchronize(c.ev1));
        float ms = 0.f;
        TRT_CHECK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
        c.sweeps += 1;
        c.sweep_ms += ms;
        trt::pk_check(c);
    });
}

int trt_rt_bc_update(trt_ctx *ctx)
{
    return guarded(ctx, [&](Ctx &c) { trt::rt_bc_update(c); TRT_CHECK(hipStreamSynchronize(c.stream)); });
}

int trt_mloqd_fv(trt_ctx *ctx, int *kits)
{
    return guarded(ctx, [&](Ctx &c) { trt::mloqd_fv(c, kits); TRT_CHECK(hipStreamSynchronize(c.stream)); });
}

int trt_egp_fv_newt(trt_ctx *ctx, int mgqd_it, int *its)
{
    return guarded(ctx, [&](Ctx &c) {
        const int n = trt::egp_fv_newt(c, mgqd_it);
        if (its) *its = n;
        TRT_CHECK(hipStreamSynchronize(c.stream));
    });
}

int trt_convergence(trt_ctx *ctx, const char *x, const char *x_old, const char *x_old2, int conv_type, double eps,
                    int its, double *norm, double *rho, int *converged)
{
    return guarded(ctx, [&](Ctx &c) {
        const size_t n = trt_array_size(ctx, x);
        if (n == 0 || n != trt_array_size(ctx, x_old) || n != trt_array_size(ctx, x_old2))
            throw Error("convergence arrays must exist and have equal sizes");
        trt::convergence(c, c.arr(x), c.arr(x_old), c.arr(x_old2), (int)n, conv_type, eps, its, norm, rho, converged);
    });
}

int trt_meb_solve(trt_ctx *ctx, const char *eg)
{
    return guarded(ctx, [&](Ctx &c) {
        if (!eg) throw Error("null argument");
        trt::meb_solve(c, eg);
        TRT_CHECK(hipStreamSynchronize(c.stream));
    });
}

int trt_step(trt_ctx *ctx, int its_out[4])
{
    return guarded(ctx, [&](Ctx &c) {
        trt::step(c, its_out);
        trt::pk_check(c);
    });
}

int trt_step_given_qdf(trt_ctx *ctx, int its_out[4])
{
    return guarded(ctx, [&](Ctx &c) { trt::step(c, its_out, 1); });
}

double trt_time(const trt_ctx *ctx) { return ctx ? ctx->c.time : 0.0; }

int trt_restart_out(trt_ctx *ctx, const char *path)
{
    return guarded(ctx, [&](Ctx &c) { trt::restart_out(c, path); });
}

int trt_restart_in(trt_ctx *ctx, const char *path)
{
    return guarded(ctx, [&](Ctx &c) { trt::restart_in(c, path); });
}

int trt_sweep_timed(trt_ctx *ctx, int nsweeps, double *ms)
{
    return guarded(ctx, [&](Ctx &c) {
        if (nsweeps < 1) throw Error("nsweeps must be positive");
        TRT_CHECK(hipEventRecord(c.ev0, c.stream));
        for (int k = 0; k < nsweeps; ++k) trt::transport_sweep(c);
        TRT_CHECK(hipEventRecord(c.ev1, c.stream));
        TRT_CHECK(hipEventSynchronize(c.ev1));
        float t = 0.f;
        TRT_CHECK(hipEventElapsedTime(&t, c.ev0, c.ev1));
        c.sweeps += nsweeps;
        c.sweep_ms += t;
        if (ms) *ms = t;
        trt::pk_check(c);
    });
}

int trt_sweep_kernel_timed(trt_ctx *ctx, int nsweeps, double *wall_ms, double *kernel_ms, int *launches)
{
    return guarded(ctx, [&](Ctx &c) {
        if (nsweeps < 1) throw Error("nsweeps must be positive");
        const int nl = trt::sweep_launches(c);
        std::vector<hipEvent_t> ev(2 * (size_t)nl);
        for (auto &e : ev) TRT_CHECK(hipEventCreate(&e));
        double kern = 0.0, wall = 0.0;
        for (int s = 0; s < nsweeps; ++s) {
            trt::transport_sweep(c, &ev);
            TRT_CHECK(hipEventSynchronize(ev.back()));
            for (int d = 0; d < nl; ++d) {
                float t = 0.f;
                TRT_CHECK(hipEventElapsedTime(&t, ev[2 * d], ev[2 * d + 1]));
                kern += t;
            }
            float w = 0.f;
            TRT_CHECK(hipEventElapsedTime(&w, ev.front(), ev.back()));
            wall += w;
        }
        for (auto &e : ev) (void)hipEventDestroy(e);
        c.sweeps += nsweeps;
        c.sweep_ms += wall;
        if (wall_ms) *wall_ms = wall;
        if (kernel_ms) *kernel_ms = kern;
        if (launches) *launches = nl * nsweeps;
        trt::pk_check(c);
    });
}

int trt_calib_copy(size_t nbytes, int reps, double *ms)
{
    try {
        trt::calib_copy(nbytes, reps, ms);
        return 0;
    } catch (const Error &e) {
        g_create_error = e.msg;
    }
    return -2;
}

int trt_sweep_stats(const trt_ctx *ctx, long long *sweeps, double *ms)
{
    if (!ctx) return -1;
    if (sweeps) *sweeps = ctx->c.sweeps;
    if (ms) *ms = ctx->c.sweep_ms;
    return 0;
}

int trt_qd_stats(const trt_ctx *ctx, long long *factorized, long long *reused, long long *refine_steps)
{
    if (!ctx) return -1;
    if (factorized) *factorized = ctx->c.nd_factor;
    if (reused) *reused = ctx->c.nd_reused;
    if (refine_steps) *refine_steps = ctx->c.nd_reuse_steps;
    return 0;
}

int trt_sweep_config(const trt_ctx *ctx, int *mode, int *tile_x, int *tile_y, int *grid)
{
    if (!ctx) return -1;
    if (mode) *mode = ctx->c.sweep_mode;
    if (tile_x) *tile_x = ctx->c.sg.TX;
    if (tile_y) *tile_y = ctx->c.sg.TY;
    if (grid) *grid = trt::pk_grid(ctx->c);
    return 0;
}

int trt_tgen_step(trt_ctx *ctx)
{
    return guarded(ctx, [&](Ctx &c) {
        trt::tgen_step(c);
        TRT_CHECK(hipStreamSynchronize(c.stream));
        trt::pk_check(c);
    });
}

int trt_batch_create(trt_batch **out, trt_ctx *const *ctxs, int n) { return trt_batch_create_seg(out, ctxs, n, 0); }

int trt_batch_create_seg(trt_batch **out, trt_ctx *const *ctxs, int n, int seg)
{
    if (!out) return -1;
    *out = nullptr;
    try {
        if (!ctxs || n < 1) throw Error("a batch needs at least one context");
        if (seg < 0) throw Error("segment length must be 0 (automatic) or positive");
        std::vector<Ctx *> v;
        for (int i = 0; i < n; ++i) {
            if (!ctxs[i]) throw Error("null context in batch");
            v.push_back(&ctxs[i]->c);
        }
        trt_batch *b = new trt_batch();
        try {
            b->b = trt::pk_batch_create(v, false, seg);
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
        return 0;
    } catch (const Error &e) {
        g_create_error = e.msg;
    } catch (const std::exception &e) {
        g_create_error = e.what();
    }
    return -2;
}

int trt_batch_destroy(trt_batch *b)
{
    if (!b) return -1;
    trt::pk_batch_destroy(b->b);
    delete b;
    return 0;
}

const char *trt_batch_last_error(const trt_batch *b) { return b ? b->err.c_str() : g_create_error.c_str(); }

}  // extern "C"

namespace {
template <typename F>
int bguarded(trt_batch *b, F &&f)
{
    if (!b || !b->b) return -1;
    try {
        f(*b->b);
        return 0;
    } catch (const Error &e) {
        b->err = e.msg;
    } catch (const std::exception &e) {
        b->err = e.what();
    }
    return -2;
}

void sync_members(trt::PkBatch &pb)
{
    for (Ctx *c : trt::pk_batch_ctxs(pb)) TRT_CHECK(hipStreamSynchronize(c->stream));
}
}  // namespace

extern "C" {

int trt_batch_sweep_timed(trt_batch *b, int nsweeps, double *ms)
{
    return bguarded(b, [&](trt::PkBatch &pb) {
        if (nsweeps < 1) throw Error("nsweeps must be positive");
        for (int k = 0; k < nsweeps; ++k) trt::pk_batch_sweep(pb, (k == 0 ? 1 : 0) | (k == nsweeps - 1 ? 2 : 0));
        const double t = trt::pk_batch_elapsed_ms(pb);
        trt::pk_batch_check(pb);
        sync_members(pb);
        for (Ctx *c : trt::pk_batch_ctxs(pb)) c->sweep_ms += t / (double)trt::pk_batch_ctxs(pb).size();
        if (ms) *ms = t;
    });
}

int trt_batch_tgen_step(trt_batch *b, double *sweep_ms)
{
    return bguarded(b, [&](trt::PkBatch &pb) {
        if (trt::pk_batch_step_graphs(pb)) {
            trt::pk_batch_tgen_step_graph(pb);
        } else {
            for (Ctx *c : trt::pk_batch_ctxs(pb)) trt::tgen_pre(*c);
            trt::pk_batch_sweep(pb, 3);
            for (Ctx *c : trt::pk_batch_ctxs(pb)) trt::tgen_post(*c);
        }
        sync_members(pb);
        const double t = trt::pk_batch_elapsed_ms(pb);
        trt::pk_batch_check(pb);
        if (sweep_ms) *sweep_ms = t;
    });
}

int trt_batch_step(trt_batch *b, int *its)
{
    return bguarded(b, [&](trt::PkBatch &pb) {
        std::vector<int> v;
        trt::batch_step(pb, its ? &v : nullptr);
        sync_members(pb);
        if (its) std::copy(v.begin(), v.end(), its);
    });
}

int trt_batch_use_graph(trt_batch *b, int on)
{
    return bguarded(b, [&](trt::PkBatch &pb) {
        if (on < 0 || on > 2) throw Error("graph mode must be 0, 1 or 2");
        trt::pk_batch_use_graph(pb, on);
    });
}

int trt_batch_info(const trt_batch *b, long long *updates_per_sweep, int *grid, int *nitems)
{
    if (!b || !b->b) return -1;
    trt::pk_batch_info(*b->b, updates_per_sweep, grid, nitems);
    return 0;
}

int trt_batch_segment_length(const trt_batch *b) { return (b && b->b) ? trt::pk_batch_seg(*b->b) : -1; }

}  // extern "C"
