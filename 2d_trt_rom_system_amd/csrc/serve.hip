// Request loop with continuous batching (no counterpart in the reference, which
// runs one problem per process).
//
// A request is a problem (trt_problem), an optional prescribed temperature
// field and a number of steps: Tgen_QDf steps (ALGORITHMS.f08:1031-1095: one sweep
// per step with the factors recomputed, no low-order solve), or in the server's solve
// mode full time steps of TRT_MLQD_ALGORITHM (a solve: sweeps, QD factors, MLOQD and
// grey Newton iterations, from the initial state).  Requests wait in arrival order.
// At each tick the loop
//   1. admits waiting requests that have arrived, in order, while the device
//      bytes of the active set stay within the HBM budget and the active count
//      within its cap (a request larger than the whole budget fails; a solve is
//      admitted with its low-order state allocated);
//   2. runs one step of every active request, all sweeps of a Tgen_QDf tick, or of
//      each round of RT iterations of a solve tick, in ONE persistent launch (the
//      batch is rebuilt only when membership changes), a solve's low-order work on
//      its own stream from a host thread of its own;
//   3. retires finished requests: their outputs are copied to host memory and
//      their device state is freed, so waiting requests can be admitted.
// Each request keeps its own device state and stream; the batched sweep does
// not change any problem's arithmetic, so a request's results are the same
// alone and in any mix.
#include "trt_internal.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>

namespace trt {

namespace {

// Outputs kept for a finished request (what Tgen_QDf writes per step).
const char *const kOutputs[] = {"Temp",      "E_avg",      "HO_E_avg",   "HO_Eg_avg",  "HO_Eg_edgV", "HO_Eg_edgH",
                                "fg_avg_xx", "fg_avg_yy",  "fg_avg_xy",  "fg_edgV_xx", "fg_edgV_xy",
                                "fg_edgH_yy", "fg_edgH_xy", "Cg_L",      "Cg_B",       "Cg_R",
                                "Cg_T"};

struct Request {
    // owned copy of the problem and its vectors
    trt_problem p{};
    std::vector<double> dx, dy, ox, oy, w, nu, kmult, temp;
    std::vector<int> mat;
    int steps = 0, done_steps = 0;
    double arrival = 0.0, admit = -1.0, done = -1.0;
    int status = 1;            // 1 waiting, 2 active, 0 done, -1 failed
    std::string error;
    size_t bytes = 0;
    long long updates = 0;
    trt_ctx *ctx = nullptr;
    // outputs, in pinned host memory (copied there while draining; kept until
    // the server is destroyed): offsets of each output in pin
    double *pin = nullptr;
    size_t pin_n = 0;
    std::vector<size_t> out_off;
    hipEvent_t drained = nullptr;
};

}  // namespace

struct Server {
    size_t budget = 0;
    int max_active = 0;
    int mode = 0;              // 0: Tgen_QDf steps; 1: full time steps (solves)
    std::vector<Request> reqs;
    // every request's device state lives in one arena of the budget's size,
    // reserved before the first run's clock starts and held until destruction
    Arena *arena = nullptr;
    double setup_s = 0.0;      // seconds spent reserving the arena
    // statistics of the last run
    double wall = 0.0;
    long long ticks = 0, updates = 0;
    size_t peak_bytes = 0;
    int peak_active = 0;
    double sweep_ms = 0.0;
    double t_admit = 0.0, t_rebuild = 0.0, t_step = 0.0, t_retire = 0.0;   // host seconds per phase
    ~Server()
    {
        for (Request &r : reqs)
            if (r.pin) (void)hipHostFree(r.pin);
        arena_destroy(arena);
    }
};

}  // namespace trt

struct trt_server {
    trt::Server s;
    std::string err;
};

namespace {

using trt::Error;
using trt::Request;
using trt::kOutputs;

template <typename T>
std::vector<T> copy_vec(const T *p, size_t n, const char *what)
{
    if (!p) throw Error(std::string("null problem vector ") + what);
    return std::vector<T>(p, p + n);
}

void rebind(Request &r)
{
    r.p.dx = r.dx.data(); r.p.dy = r.dy.data();
    r.p.ox = r.ox.data(); r.p.oy = r.oy.data(); r.p.w = r.w.data();
    r.p.nu = r.nu.data(); r.p.kmult = r.kmult.data(); r.p.mat = r.mat.data();
}

// admission estimate: the creation bytes plus room for block rounding and the
// sweep's work list.  Creation itself fails when the arena has no room.
size_t admission_bytes(const trt_problem &p)
{
    const size_t b = trt_ctx_bytes(&p);
    return b + b / 16 + ((size_t)8 << 20);
}

// Retirement, in two halves so the copies overlap the next ticks: the outputs
// are copied to a pinned buffer on the request's stream (drain_start); once the
// copy has completed they move to host vectors and the request's device blocks
// go back to the block cache (drain_finish).
void drain_start(trt::Server &s, Request &r)
{
    trt::Ctx &c = r.ctx->c;
    const size_t nout = sizeof(kOutputs) / sizeof(kOutputs[0]);
    r.out_off.assign(nout + 1, 0);
    for (size_t k = 0; k < nout; ++k) r.out_off[k + 1] = r.out_off[k] + trt_array_size(r.ctx, kOutputs[k]);
    r.pin_n = r.out_off[nout];
    TRT_CHECK(hipHostMalloc(reinterpret_cast<void **>(&r.pin), r.pin_n * sizeof(double), hipHostMallocDefault));
    for (size_t k = 0; k < nout; ++k)
        TRT_CHECK(hipMemcpyAsync(r.pin + r.out_off[k], c.arr(kOutputs[k]),
                                 (r.out_off[k + 1] - r.out_off[k]) * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    if (!r.drained) TRT_CHECK(hipEventCreateWithFlags(&r.drained, hipEventDisableTiming));
    TRT_CHECK(hipEventRecord(r.drained, c.stream));
    r.status = 3;
}

// true when the request's copy had completed (or wait) and it is now done
bool drain_finish(Request &r, bool wait, double now)
{
    if (wait) {
        TRT_CHECK(hipEventSynchronize(r.drained));
    } else {
        const hipError_t e = hipEventQuery(r.drained);
        if (e == hipErrorNotReady) return false;
        TRT_CHECK(e);
    }
    (void)hipEventDestroy(r.drained);
    r.drained = nullptr;
    trt::ctx_destroy(r.ctx);
    r.ctx = nullptr;
    r.status = 0;
    r.done = now;
    return true;
}

void fail(Request &r, const std::string &msg, double now)
{
    if (r.drained) {
        (void)hipEventSynchronize(r.drained);
        (void)hipEventDestroy(r.drained);
        r.drained = nullptr;
    }
    if (r.ctx) trt::ctx_destroy(r.ctx);
    r.ctx = nullptr;
    r.status = -1;
    r.error = msg;
    r.done = now;
}

void run(trt::Server &s)
{
    using clock = std::chrono::steady_clock;
    if (!s.arena) {
        // a server reserves its budget once; the time it takes is not the requests'
        const auto ta = clock::now();
        s.arena = trt::arena_create(s.budget);
        s.setup_s = std::chrono::duration<double>(clock::now() - ta).count();
    }
    const auto t0 = clock::now();
    auto now = [&]() { return std::chrono::duration<double>(clock::now() - t0).count(); };
    // the waiting requests in arrival order (finished or failed ones of earlier runs stay out)
    std::vector<size_t> order;
    for (size_t i = 0; i < s.reqs.size(); ++i)
        if (s.reqs[i].status == 1) order.push_back(i);
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return s.reqs[a].arrival < s.reqs[b].arrival; });
    size_t next = 0;                       // first request in arrival order not yet admitted
    s.ticks = 0; s.updates = 0; s.peak_bytes = 0; s.peak_active = 0; s.sweep_ms = 0.0;
    s.t_admit = s.t_rebuild = s.t_step = s.t_retire = 0.0;
    std::vector<size_t> active, draining;
    trt::PkBatch *batch = nullptr;
    bool changed = true;
    trt::Arena *arena = s.arena;
    const char *venv = getenv("TRT_VERBOSE");
    const bool trace = venv && atoi(venv) > 0;
    auto active_bytes = [&]() { return trt::arena_used(arena); };

    // admits arrived requests in order while they fit; their initialisation
    // runs on their own streams
    auto admit = [&]() {
        while (next < order.size()) {
            Request &r = s.reqs[order[next]];
            if (r.arrival > now()) break;
            const size_t need = admission_bytes(r.p);
            if (need > s.budget) {
                fail(r, "request needs more device memory than the whole budget", now());
                ++next;
                continue;
            }
            if ((int)active.size() >= s.max_active || active_bytes() + need > s.budget) break;
            const double ta = now();
            trt_ctx *ctx = nullptr;
            try {
                if (s.mode == 1 && !r.temp.empty())
                    throw Error("a solve starts from the initial state (TEMP_INIT): no temperature field");
                ctx = trt::ctx_create(&r.p, true, arena);
                if (s.mode == 1) trt::prepare_solves(ctx->c);
            } catch (const Error &e) {
                if (ctx) trt::ctx_destroy(ctx);
                ctx = nullptr;
                const bool nomem = e.msg.find("arena exhausted") != std::string::npos;
                if (nomem && !(active.empty() && draining.empty())) break;   // wait for memory
                fail(r, nomem ? std::string("request needs more device memory than the whole budget")
                              : "context creation: " + e.msg, now());
                ++next;
                continue;
            }
            r.ctx = ctx;
            if (!r.temp.empty() && trt_push(ctx, "Temp", r.temp.data(), r.temp.size()) != 0)
                throw Error("push Temp: " + ctx->c.err);
            r.bytes = ctx->c.bytes;
            r.admit = now();
            s.t_admit += r.admit - ta;
            r.status = 2;
            active.push_back(order[next]);
            ++next;
            changed = true;
        }
    };
    // completes the drains whose copies have finished (all of them with wait)
    auto drain = [&](bool wait) {
        const double tr = now();
        std::vector<size_t> still;
        for (size_t i : draining)
            if (!drain_finish(s.reqs[i], wait, now())) still.push_back(i);
        draining.swap(still);
        s.t_retire += now() - tr;
    };

    try {
        while (next < order.size() || !active.empty() || !draining.empty()) {
            drain(false);
            admit();
            if (active.empty()) {
                if (!draining.empty()) {
                    drain(true);
                } else if (next < order.size()) {
                    const double wait = s.reqs[order[next]].arrival - now();
                    if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.05)));
                }
                continue;
            }
            s.peak_active = std::max(s.peak_active, (int)active.size());
            s.peak_bytes = std::max(s.peak_bytes, active_bytes());
            // one Tgen_QDf step of every active request, the sweeps batched
            if (changed || !batch) {
                const double tb = now();
                trt::pk_batch_destroy(batch);
                batch = nullptr;
                std::vector<trt::Ctx *> cs;
                for (size_t i : active) cs.push_back(&s.reqs[i].ctx->c);
                batch = trt::pk_batch_create(cs, true);
                changed = false;
                s.t_rebuild += now() - tb;
            }
            const double ts = now();
            const std::vector<size_t> ticked = active;
            std::vector<int> its;
            if (s.mode == 1) {
                trt::batch_step(*batch, &its);
            } else {
                for (size_t i : ticked) trt::tgen_pre(s.reqs[i].ctx->c);
                trt::pk_batch_sweep(*batch, 3);
                for (size_t i : ticked) trt::tgen_post(s.reqs[i].ctx->c);
            }
            // host work while the tick runs: arrivals join the next tick, finished copies complete
            const double th = now();
            admit();
            drain(false);
            const double host_during = now() - th;
            for (size_t i : ticked) TRT_CHECK(hipStreamSynchronize(s.reqs[i].ctx->c.stream));
            s.sweep_ms += trt::pk_batch_elapsed_ms(*batch);
            trt::pk_batch_check(*batch);
            ++s.ticks;
            s.t_step += now() - ts - host_during;
            if (trace)
                fprintf(stderr, "[serve] t=%.3f tick %lld: %zu active, %zu draining, %zu waiting, tick %.1f ms (host %.1f ms), arena %.1f GiB\n",
                        now(), s.ticks, ticked.size(), draining.size(), order.size() - next, 1e3 * (now() - ts),
                        1e3 * host_during, arena_used(arena) / 1073741824.0);
            s.peak_bytes = std::max(s.peak_bytes, active_bytes());
            // requests that finished start draining; the rest stay active
            std::vector<size_t> still;
            for (size_t i : active) {
                Request &r = s.reqs[i];
                const auto pos = std::find(ticked.begin(), ticked.end(), i);
                if (pos == ticked.end()) {
                    still.push_back(i);        // admitted during this tick
                    continue;
                }
                const trt::Ctx &c = r.ctx->c;
                r.bytes = c.bytes;
                r.done_steps += 1;
                const long long sweeps = s.mode == 1 ? its[4 * (size_t)(pos - ticked.begin()) + 3] : 1;
                const long long u = (long long)c.nx * c.ny * c.nm * c.ng * sweeps;
                r.updates += u;
                s.updates += u;
                if (r.done_steps >= r.steps) {
                    const double tr = now();
                    drain_start(s, r);
                    s.t_retire += now() - tr;
                    draining.push_back(i);
                    changed = true;
                } else {
                    still.push_back(i);
                }
            }
            active.swap(still);
            if (changed) {
                trt::pk_batch_destroy(batch);
                batch = nullptr;
            }
        }
    } catch (...) {
        trt::pk_batch_destroy(batch);
        for (size_t i : active) fail(s.reqs[i], "aborted by an error of the request loop", now());
        for (size_t i : draining) fail(s.reqs[i], "aborted by an error of the request loop", now());
        s.wall = now();
        trt::cache_trim();
        throw;
    }
    trt::pk_batch_destroy(batch);
    s.wall = now();
    trt::cache_trim();
}

template <typename F>
int sguarded(trt_server *sv, F &&f)
{
    if (!sv) return -1;
    try {
        f(sv->s);
        return 0;
    } catch (const Error &e) {
        sv->err = e.msg;
    } catch (const std::exception &e) {
        sv->err = e.what();
    }
    return -2;
}

}  // namespace

extern "C" {

int trt_server_create(trt_server **out, size_t hbm_budget, int max_active)
{
    if (!out) return -1;
    *out = nullptr;
    if (hbm_budget == 0 || max_active < 1) return -1;
    trt_server *sv = new trt_server();
    sv->s.budget = hbm_budget;
    sv->s.max_active = max_active;
    *out = sv;
    return 0;
}

int trt_server_destroy(trt_server *sv)
{
    if (!sv) return -1;
    for (Request &r : sv->s.reqs)
        if (r.ctx) trt::ctx_destroy(r.ctx);
    delete sv;
    return 0;
}

const char *trt_server_last_error(const trt_server *sv) { return sv ? sv->err.c_str() : ""; }

int trt_server_set_mode(trt_server *sv, int mode)
{
    return sguarded(sv, [&](trt::Server &s) {
        if (mode != 0 && mode != 1) throw Error("server mode must be 0 (Tgen_QDf steps) or 1 (full time steps)");
        s.mode = mode;
    });
}

int trt_server_submit(trt_server *sv, const trt_problem *p, const double *temp, int steps, double arrival_s,
                      int *id)
{
    return sguarded(sv, [&](trt::Server &s) {
        if (!p) throw Error("null problem");
        if (steps < 1) throw Error("steps must be positive");
        if (p->sweep_mode == 1) throw Error("served requests use the tile engine (sweep_mode 0, 2 or 3)");
        Request r;
        r.p = *p;
        const size_t na = (size_t)p->nx * p->ny;
        if (p->nx < 2 || p->ny < 2 || p->nm < 4 || p->ng < 1 || p->nmat < 1) throw Error("invalid problem sizes");
        r.dx = copy_vec(p->dx, p->nx, "dx");
        r.dy = copy_vec(p->dy, p->ny, "dy");
        r.ox = copy_vec(p->ox, p->nm, "ox");
        r.oy = copy_vec(p->oy, p->nm, "oy");
        r.w = copy_vec(p->w, p->nm, "w");
        r.nu = copy_vec(p->nu, (size_t)p->ng + 1, "nu");
        r.kmult = copy_vec(p->kmult, p->nmat, "kmult");
        r.mat = copy_vec(p->mat, na, "mat");
        if (temp) r.temp.assign(temp, temp + na);
        r.steps = steps;
        r.arrival = arrival_s;
        s.reqs.push_back(std::move(r));
        rebind(s.reqs.back());
        if (id) *id = (int)s.reqs.size() - 1;
    });
}

int trt_server_run(trt_server *sv)
{
    return sguarded(sv, [&](trt::Server &s) {
        for (Request &r : s.reqs) rebind(r);
        run(s);
    });
}

const char *trt_server_request_error(const trt_server *sv, int id)
{
    if (!sv || id < 0 || id >= (int)sv->s.reqs.size()) return "";
    return sv->s.reqs[id].error.c_str();
}

int trt_server_request(const trt_server *sv, int id, trt_request_stats *out)
{
    if (!sv || !out || id < 0 || id >= (int)sv->s.reqs.size()) return -1;
    const Request &r = sv->s.reqs[id];
    out->arrival_s = r.arrival;
    out->admit_s = r.admit;
    out->done_s = r.done;
    out->updates = r.updates;
    out->device_bytes = r.bytes;
    out->steps = r.done_steps;
    out->status = r.status;
    return 0;
}

int trt_server_summary(const trt_server *sv, trt_server_stats *out)
{
    if (!sv || !out) return -1;
    const trt::Server &s = sv->s;
    out->wall_s = s.wall;
    out->ticks = s.ticks;
    out->updates = s.updates;
    out->sweep_ms = s.sweep_ms;
    out->peak_bytes = s.peak_bytes;
    out->peak_active = s.peak_active;
    out->admit_s = s.t_admit;
    out->rebuild_s = s.t_rebuild;
    out->step_s = s.t_step;
    out->retire_s = s.t_retire;
    out->setup_s = s.setup_s;
    return 0;
}

size_t trt_server_output_size(const trt_server *sv, int id, const char *name)
{
    if (!sv || !name || id < 0 || id >= (int)sv->s.reqs.size()) return 0;
    const Request &r = sv->s.reqs[id];
    if (r.status != 0) return 0;
    for (size_t k = 0; k < sizeof(kOutputs) / sizeof(kOutputs[0]); ++k)
        if (!strcmp(kOutputs[k], name)) return r.out_off[k + 1] - r.out_off[k];
    return 0;
}

int trt_server_output(const trt_server *sv, int id, const char *name, double *host, size_t n)
{
    const size_t m = trt_server_output_size(sv, id, name);
    if (m == 0 || !host || n != m) return -1;
    const Request &r = sv->s.reqs[id];
    for (size_t k = 0; k < sizeof(kOutputs) / sizeof(kOutputs[0]); ++k)
        if (!strcmp(kOutputs[k], name)) std::memcpy(host, r.pin + r.out_off[k], m * sizeof(double));
    return 0;
}

}  // extern "C"
