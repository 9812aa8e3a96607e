// Internal declarations shared by the HIP translation units of the TRT library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/trt_capi.h"

#define TRT_CHECK(expr)                                                                    \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess) {                                                            \
            throw trt::Error(std::string(#expr) + ": " + hipGetErrorString(_e));          \
        }                                                                                  \
    } while (0)

namespace trt {

// Intensities stored in FP32 carry a factor 2^40: the cold high groups (near
// 1e-40) then stay clear of FP32's denormal range in the FP32 corner solve,
// whose largest intermediates (coefficient^3 x intensity) stay below 1e36.  The
// factor enters with the initial values, the FP32 sources and the boundary
// intensities, and leaves in the FP64 moments and at push/pull; a power of two,
// it changes no rounding.
constexpr double kIScale32 = 1099511627776.0;   // 2^40
template <typename TI>
__host__ __device__ constexpr double iscale() { return sizeof(TI) == 4 ? kIScale32 : 1.0; }

// Lanes of a face-exchange or boundary slice: the lane count rounded up to 8, so that each
// (tile, chunk) run of 16-byte chunks fills whole 128-byte cache lines and no line holds
// records of two tiles (written by different workgroups at different times).
__host__ __device__ inline int rec_lanes(int NL) { return (NL + 7) & ~7; }

// Element index of value t of the record of (tile, lane) in a face-exchange or boundary
// slice whose records hold N values of TI: records are stored in 16-byte chunks,
// chunk-major ([tile][chunk][lane][16 bytes]).
template <typename TI>
__host__ __device__ inline size_t rec_elem(int tile, int t, int N, int NL, int lane)
{
    constexpr int E = 16 / (int)sizeof(TI);
    return ((size_t)(tile * (N / E) + t / E) * NL + lane) * E + t % E;
}

struct Error {
    std::string msg;
    explicit Error(std::string m) : msg(std::move(m)) {}
};

constexpr double PI = 3.14159265358979323846264338327950288;
constexpr double H_PLANCK = 6.62613e-19;   // erg*sh (MainSrc/main.f08:18)
constexpr double ERG = 6.24150934e11;      // eV/erg (MainSrc/main.f08:21)
constexpr double C_CONST = 2.99792458e2;   // cm/sh, as hard-coded in several reference routines

// A device array registered under its reference name with its Fortran shape.
struct DevArray {
    double *p = nullptr;
    int d[4] = {1, 1, 1, 1};
    size_t size() const { return (size_t)d[0] * d[1] * d[2] * d[3]; }
};

// Sweep launch geometry.
struct SweepGeom {
    int K = 0;        // directions per quadrant (N_m/4)
    int gblk = 0;     // groups per block (gblk*K <= 256)
    int nblk_g = 0;   // ceil(ng/gblk)
    int TX = 8, TY = 8, NTX = 0, NTY = 0;
};

// Quadrant partial moments of a sweep: 13 FP64 values per (cell, quadrant, group), the
// direction sums of COLLAPSE_INTENSITIES (TRANSPORT_SOLVES.f08:447-470) over one
// quadrant: the cell average's w Ox^2, w Oy^2, w Ox Oy, w (values 0-3), the downwind V
// face's w Ox^2, w Ox Oy, w Ox, w (4-7) and the downwind H face's w Oy^2, w Ox Oy, w Ox,
// w Oy, w (8-12), each times (I + eps).  They are stored by work item of the tile engine:
// item (quadrant q, group block gb, tile (txl, tyl) of q's swept frame) holds NMOM runs
// of PP values, run v holding value v of the item's (cell, group) pairs p = (r TX + ii)
// gblk + gl (r, ii: row and column in the tile; gl: group in the block), so one store
// instruction of the moment reduction writes 64 consecutive pairs.  Incoming boundary
// faces (no upwind cell in the quadrant) have arrays of their own: V [q][j][g][4], H
// [q][i][g][5].
constexpr int NMOM = 13;

struct PartLayout {
    int nx = 0, ny = 0, ng = 0, TX = 1, TY = 1, NTX = 0, NTY = 0, gblk = 1, nblk_g = 0, PP = 0;
    // first value of run 0 of an item
    __host__ __device__ __forceinline__ size_t item(int q, int gb, int txl, int tyl) const
    {
        return ((((size_t)q * nblk_g + gb) * NTY + tyl) * NTX + txl) * NMOM * (size_t)PP;
    }
    // value v of cell (i, j), group g, quadrant q
    __host__ __device__ __forceinline__ size_t at(int i, int j, int q, int g, int v) const
    {
        const bool px = (q == 0 || q == 3), py = q < 2;
        const int il = px ? i : nx - 1 - i, jl = py ? j : ny - 1 - j;
        const int txl = il / TX, tyl = jl / TY, gb = g / gblk;
        const int p = ((jl - tyl * TY) * TX + (il - txl * TX)) * gblk + (g - gb * gblk);
        return item(q, gb, txl, tyl) + (size_t)v * PP + p;
    }
    size_t size() const { return (size_t)4 * nblk_g * NTY * NTX * NMOM * PP; }
};

// value v (0-3) of V face f of row j, quadrant q: the downwind V face of the cell upwind of
// it, or the incoming boundary face
__host__ __device__ __forceinline__ double part_v(const PartLayout &L, const double *Part, const double *PVb, int f,
                                                  int j, int q, int g, int v)
{
    const int i = (q == 0 || q == 3) ? f - 1 : f;
    return (i < 0 || i >= L.nx) ? PVb[(((size_t)q * L.ny + j) * L.ng + g) * 4 + v] : Part[L.at(i, j, q, g, 4 + v)];
}

// value v (0-4) of H face f of column i, quadrant q
__host__ __device__ __forceinline__ double part_h(const PartLayout &L, const double *Part, const double *PHb, int i,
                                                  int f, int q, int g, int v)
{
    const int j = q < 2 ? f - 1 : f;
    return (j < 0 || j >= L.ny) ? PHb[(((size_t)q * L.nx + i) * L.ng + g) * 5 + v] : Part[L.at(i, j, q, g, 8 + v)];
}

// Batched reduced-QD systems (QD_FV) stored as coefficient arrays on device.
struct NDPlan;      // nested-dissection plan and fronts of the direct QD solver (qd_nd.hip)

struct QDSystems {
    int nsys = 0, nx = 0, ny = 0;
    size_t nrow = 0;              // unknowns per system: 3 nx ny + nx + ny
    double *coef = nullptr;       // [nsys][16][nx*ny]
    double *bnd = nullptr;        // [nsys][8][max(nx,ny)]
    double *diag = nullptr;       // [nsys][nrow] Jacobi diagonal of A S
    double *cs = nullptr;         // [nsys][nrow] column scales S
    double *ref = nullptr;        // [nsys][nrow] scale reference (when not the initial guess)
    double *x0 = nullptr;         // [nsys][nrow] initial guess, kept for restarts
    double *b = nullptr;          // [nsys][nrow]
    double *x = nullptr;          // [nsys][nrow] solution (initial guess on entry)
    double *work = nullptr;       // 7 * nsys * nrow
    double *red = nullptr;        // reduction partials
    double *scal = nullptr;       // per-system scalars
    int *flags = nullptr;         // per-system convergence flags (device)
    int red_blocks = 0;
    NDPlan *nd = nullptr;         // nested-dissection plan, fronts and factors of these systems
    // per system: how many coming solves go straight to the direct solver (set
    // when BiCGSTAB failed on it; the systems of successive iterations are close)
    std::vector<int> direct_next;
};

// Iteration history of a time step, as the reference writes it to its output file
// (MainSrc/ALGORITHMS.f08:761-829): per MLOQD iteration the Newton count, the Krylov
// count of each Newton solve and the MLOQD convergence norms; per RT iteration the MLOQD
// count, the group solves' Krylov counts of its last MLOQD_FV and the RT norms.
struct MgItLog {
    int newt = 0;
    std::vector<int> gkits;
    double tnorm = 0, enorm = 0, trho = 0, erho = 0;
};
struct RtItLog {
    int mg = 0;
    std::vector<int> kits;
    double tnorm = 0, enorm = 0, trho = 0, erho = 0;
    std::vector<MgItLog> mgl;
};

struct PkState;   // persistent-sweep work list and flags (sweep_pk.hip)
struct Arena;     // one device allocation carved into contexts' blocks (state.hip)
struct DirectWork;  // block-LU workspace of the direct QD fallback (qd_direct.hip)

struct Ctx {
    trt_problem p{};
    int nx = 0, ny = 0, nm = 0, ng = 0, nmat = 0;
    int precision = 0;
    double time = 0.0;
    std::string err;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    long long sweeps = 0;
    double sweep_ms = 0.0;

    // registry of named FP64 device arrays (reference names)
    std::map<std::string, DevArray> arrays;
    std::vector<void *> owned;       // every device block of this context
    std::vector<size_t> owned_bytes; // and its size
    bool pooled = false;             // stream and pinned buffer from the process caches
    Arena *arena = nullptr;          // when set, device blocks come from (and return to) this arena

    // problem vectors on device
    double *dx = nullptr, *dy = nullptr, *A = nullptr, *ox = nullptr, *oy = nullptr, *w = nullptr;
    double *nu = nullptr, *kmult = nullptr;
    int *mat = nullptr;
    double *bcin = nullptr;           // [4][ng] incoming boundary intensities (L,B,R,T)
    std::vector<double> h_bcin;       // host copy

    // sweep state
    SweepGeom sg;
    void *Icrn[2] = {nullptr, nullptr};   // [cell][q][L][4], FP64 or FP32
    int icur = 0;                          // buffer holding the latest I_crn
    int iold = 0;                          // buffer holding I_crn_old
    void *XV = nullptr, *XH = nullptr;     // tile face exchange
    // reflective sides (BC_Type 1): bit 0 left, 1 bottom, 2 right, 3 top; the boundary
    // slices hold the corner intensities of the domain faces, in the exchange record
    // layout, per side and quadrant: BV [left, right][q][NTY tiles][NLr lanes][2 TY],
    // BH [bottom, top][q][NTX tiles][NLr lanes][2 TX] (NLr = rec_lanes(NL)).  A side's incoming quadrants
    // read their inflow there; its outgoing quadrants write their outflow there.
    int refl = 0;
    void *BV = nullptr, *BH = nullptr;
    double *sc = nullptr;                  // [cell][g][2] = (KapE + 1/(c dt), RT_Src)
    float *scf = nullptr;                  // FP32 copy of sc for the FP32 sweep
    double *Part = nullptr;                // quadrant partial moments by work item (PartLayout)
    double *PVb = nullptr, *PHb = nullptr; // their incoming boundary faces
    PartLayout pl;
    size_t icrn_bytes = 0;
    int sweep_mode = 2;                    // 1, 2, 3 as trt_problem.sweep_mode
    PkState *pk = nullptr;

    // low-order systems
    QDSystems mg, grey;
    double *scratch = nullptr;             // general FP64 scratch
    size_t scratch_n = 0;
    double *hbuf = nullptr;                // pinned host staging for scalars
    int lin_maxit = 0;
    int verbose = 0;                       // TRT_VERBOSE=1: per-solve Krylov statistics on stderr
    double lin_rtol = 1e-13;
    std::vector<int> last_kits;
    std::vector<int> newt_kits;        // Krylov count of each solve of the latest EGP_FV_NEWT
    std::vector<RtItLog> itlog;        // iteration history of the latest time step
    DirectWork *direct = nullptr;
    int qd_solver = 0;            // 0: nested dissection, factors reused within a time step; 1: BiCGSTAB with
                                  // the direct fallback; 2: block LU; 3: nested dissection, factorizing every solve
    long long nd_factor = 0, nd_reused = 0, nd_reuse_steps = 0;   // direct-solve statistics
    // POD and DMD bases created on this context: they use its stream and arrays, so
    // the context frees any still open when it is destroyed
    std::vector<trt_pod *> pods;
    std::vector<trt_dmd *> dmds;
    std::vector<trt_romgrid *> romgrids;   // data grids of the ROM runs (csrc/romgrid.hip)

    double *arr(const char *name);
    DevArray &reg(const std::string &name, int d0, int d1 = 1, int d2 = 1, int d3 = 1);
    void *dalloc(size_t bytes);
    void dfree_last();            // frees the block of the latest dalloc
    size_t bytes = 0;             // device bytes allocated by dalloc
};

}  // namespace trt

// the C ABI's context handle
struct trt_ctx {
    trt::Ctx c;
};

namespace trt {

// --- launchers (each enqueues on ctx.stream) ---
void init_state(Ctx &c);
void material_update(Ctx &c);
void transport_sweep(Ctx &c, std::vector<hipEvent_t> *ev = nullptr);
int sweep_launches(const Ctx &c);
size_t derived_size(const Ctx &c, const char *name);
void pull_derived(Ctx &c, const char *name, double *host);
void rebuild_sc(Ctx &c);
void calib_copy(size_t nbytes, int reps, double *ms);
void collapse_intensities(Ctx &c);
// COLLAPSE_INTENSITIES followed by fg_Calc, in one pass over the partial moments
void collapse_with_fg(Ctx &c);
void fg_calc(Ctx &c);
void cg_calc(Ctx &c);
void mgqd_in_calc(Ctx &c);
void gqd_in_calc(Ctx &c);
void old_mgqd_coefs(Ctx &c);
void old_grey_coefs(Ctx &c);
void mloqd_fv(Ctx &c, int *kits);
void collapse_mg_ef(Ctx &c);
void cbar_calc(Ctx &c);
void grey_coefs(Ctx &c);
int egp_fv_newt(Ctx &c, int mgqd_it);
// MEB_SOLVE (use_grey = 0) from the group energies in the named array
void meb_solve(Ctx &c, const char *eg);
// CONVERGENCE's one-pass reduction: partial blocks for n values, and the launch that
// writes 4 x blocks partials (max |x - xo|, max |xo - xo2|, max |x| or max |1 - xo/x|,
// min |x|) to part on stream s
int conv_blocks(size_t n);
void conv_partial_launch(const double *x, const double *xo, const double *xo2, int n, int conv_type, double *part,
                         int blocks, hipStream_t s);
// the message trt_last_error(NULL) returns on this thread
void set_thread_error(const std::string &msg);
void convergence(Ctx &c, const double *x, const double *xo, const double *xo2, int n, int conv_type,
                 double eps, int its, double *norm, double *rho, int *conv);
// ho_form 0: full-order step; 1: the ROM runs' step (QD factors and C_g from the context, no sweep)
int step(Ctx &c, int its_out[4], int ho_form = 0);
// the same step in phases (loqd.hip): a batch interleaves them to sweep many problems at once
struct StepState {
    int ho_form = 0, maxit_rte = 0;
    int rt_its = 0, mg_total = 0, newt_total = 0, sweeps = 0;
    bool rt_conv = false;
};
void step_begin(Ctx &c, StepState &st, int ho_form);
bool step_done(const StepState &st);
bool step_sweeps_next(const StepState &st);
void step_rt_iteration(Ctx &c, StepState &st, bool swept);
int step_end(Ctx &c, const StepState &st, int its_out[4]);
// allocates the low-order state (QD systems, scratch, nested-dissection plans) of full steps
void prepare_solves(Ctx &c);
// full steps of every member of a batch, the sweeps of each round of RT iterations in one
// launch; its (may be null) receives RT, MLOQD, Newton and sweep counts per member
struct PkBatch;
void batch_step(PkBatch &b, std::vector<int> *its);
void pull_icrn(Ctx &c, int which, double *host);
void push_icrn(Ctx &c, int which, const double *host);
// RESTART_OUT / RESTART_IN of the context's state (csrc/restart.hip); restart_in
// checks the whole file before it changes any device array
void restart_out(Ctx &c, const char *path);
void restart_in(Ctx &c, const char *path);
// C_g of the ROM runs: BCg_next (per group: left N_y, bottom N_x, right N_y, top N_x
// values) split into Cg_L/B/R/T_next; a no-op for any other array name
void bcg_split_if(Ctx &c, const std::string &name);
// frees the POD/DMD bases still registered on the context (csrc/rom.hip)
void rom_release_all(Ctx &c);
// frees the ROM data grids still registered on the context (csrc/romgrid.hip)
void romgrid_release_all(Ctx &c);
void sweep_setup(Ctx &c);
// persistent sweep (sweep_pk.hip)
bool pk_supported(const Ctx &c);
void pk_sweep(Ctx &c, std::vector<hipEvent_t> *ev = nullptr);
void pk_check(Ctx &c);
void diag_sweep(Ctx &c, std::vector<hipEvent_t> *ev = nullptr);
void pk_destroy(Ctx &c);
int pk_grid(const Ctx &c);
// RT_BC_UPDATE (UPDATES.f08:12-105) on the boundary slices of the reflective sides: each
// incoming quadrant takes the outflow of its mirror quadrant from the last sweep
void rt_bc_update(Ctx &c);
// the inflow of every reflective side back to Bg_planck(bcT) (RT_INIT's RT_BC_UPDATE with
// BC_Type 0, INITIALIZERS.f08:106)
void reflect_reset(Ctx &c);
// batched persistent sweep over many contexts (one precision and tile shape)
struct PkBatch;
// seg: column-segment length of the work items (0 = chosen from the parallelism)
PkBatch *pk_batch_create(const std::vector<Ctx *> &ctxs, bool pooled = false, int seg = 0);
int pk_batch_seg(const PkBatch &b);
void pk_batch_destroy(PkBatch *b);
// sweep (may be null: every member): the members swept this time, the others skipped
void pk_batch_sweep(PkBatch &b, int mark, const std::vector<char> *sweep = nullptr);
// mode 0: eager; 1: the batched sweep as a graph; 2: whole Tgen_QDf steps as graphs
void pk_batch_use_graph(PkBatch &b, int mode);
bool pk_batch_step_graphs(const PkBatch &b);
void pk_batch_tgen_step_graph(PkBatch &b);
void pk_batch_check(PkBatch &b);
double pk_batch_elapsed_ms(PkBatch &b);
void pk_batch_info(const PkBatch &b, long long *updates, int *grid, int *nitems);
const std::vector<Ctx *> &pk_batch_ctxs(const PkBatch &b);
// one step of the sweep-only entry Tgen_QDf with Temp prescribed: the part
// before the sweep, the part after it, and the whole step
void tgen_pre(Ctx &c);
void tgen_post(Ctx &c);
void tgen_step(Ctx &c);
// device bytes a context allocates at creation (sweep and state arrays)
size_t state_bytes(const trt_problem &p);
// Process-wide cache of device blocks for contexts created and destroyed at a
// high rate (the request loop): freed blocks are kept and reused for requests
// of a similar size instead of going back through hipFree/hipMalloc.
void *cache_alloc(size_t &bytes);          // rounds bytes up to the block actually used
void cache_release(void *p, size_t bytes);
void cache_trim();                         // returns every cached block, buffer and stream
size_t cache_bytes();
// pinned host buffers and streams for pooled contexts: hipHostFree and hipFree
// wait for all work on the device, hipStreamCreate takes milliseconds
double *cache_pinned_get(size_t n);        // n doubles
void cache_pinned_put(double *p, size_t n);
hipStream_t cache_stream_get();
void cache_stream_put(hipStream_t s);
// A device arena: one hipMalloc of a fixed size, carved into blocks (best fit,
// coalescing frees), so contexts can be created and destroyed at a high rate
// without hipMalloc/hipFree (hipFree waits for all work on the device).
Arena *arena_create(size_t bytes);
void arena_destroy(Arena *a);
void *arena_alloc(Arena *a, size_t &bytes);   // nullptr when no free range fits; bytes rounded up
void arena_release(Arena *a, void *p, size_t bytes);
size_t arena_used(const Arena *a);
size_t arena_size(const Arena *a);
// context creation behind trt_ctx_create; pooled contexts take their stream and
// pinned buffer from the process caches and, with an arena, their device blocks
trt_ctx *ctx_create(const trt_problem *p, bool pooled, Arena *arena = nullptr);
void ctx_destroy(trt_ctx *ctx);
size_t sweep_bytes(const trt_problem &p, SweepGeom *out);
void qd_alloc(Ctx &c, QDSystems &s, int nsys);
// solve the batched reduced QD systems stored in s (coef, bnd filled); x holds the guess.
// The unknowns are scaled by |s.ref| (ref_given) or by |initial guess|.
void qd_solve(Ctx &c, QDSystems &s, int *its_out, bool ref_given);
// direct block-LU solve of the listed systems of s (their solutions replace s.x)
void qd_direct(Ctx &c, QDSystems &s, const std::vector<int> &which);
void direct_destroy(Ctx &c);
// nested-dissection direct solve of the listed systems of s (solutions replace s.x)
void qd_nd(Ctx &c, QDSystems &s, const std::vector<int> &which);
// the listed systems by iterative refinement with the factors of their last qd_nd:
// refinement steps taken, or -1 (factors absent or no longer contracting; s.x undefined)
int qd_nd_reuse(Ctx &c, QDSystems &s, const std::vector<int> &which, int max_steps);
// no reuse of the factors s holds now (their next solve factorizes afresh)
void qd_nd_forget(QDSystems &s);
// the plan (tree, fronts, factor storage) of s, created now if it does not exist
void qd_nd_prepare(Ctx &c, QDSystems &s);
void nd_destroy(Ctx &c);
// device copy helper
void dcopy(Ctx &c, double *dst, const double *src, size_t n);
void dfill(Ctx &c, double *dst, double v, size_t n);
double dmaxabs(Ctx &c, const double *x, size_t n);

}  // namespace trt
