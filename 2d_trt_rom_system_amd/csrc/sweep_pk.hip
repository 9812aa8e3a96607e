// K1 (tile engine): the SCB sweep of TRANSPORT_SCB (MainSrc/TRANSPORT_SOLVES.f08:14-327)
// with the angular sums of COLLAPSE_INTENSITIES (:381-511), in two launch forms
// that share one tile body:
//
//  * persistent (sweep_mode 2): ONE launch per sweep.  Work items are
//    (quadrant, group block, tile); an item depends on the tiles left of and
//    below it in its quadrant's reflected frame.  Items are listed by tile
//    diagonal, so the list is a topological order, and resident workgroups
//    take them from an atomic counter.  Every item a workgroup waits for was
//    taken earlier by a running workgroup, so the grid size never matters for
//    progress.  A workgroup waits (bounded spin) for the two upwind tiles'
//    completion flags, which carry the sweep's epoch and are never cleared.
//  * per-diagonal (sweep_mode 3): one launch per tile diagonal; kernel
//    boundaries order the tiles.
//  * batched: the persistent form over the items of many problems (uneven
//    grids, group counts and quadratures, one tile shape), merged by tile
//    diagonal so the merged list is still a topological order.  A small
//    problem's fill and drain then overlap the others' work.
//
// Tile body.  A thread owns one (quadrant, group, direction).  At the start of
// a tile it issues every load the tile needs -- the old corner intensities and
// sweep coefficients of all TX x TY cells and the incoming faces -- so a tile
// costs one memory round trip, then solves the cells in upwind order with the
// corner-edge intensities in registers.  The corner averages of every (cell,
// direction) are staged in LDS; after one block barrier wave w of the block
// sums a quarter of the directions for every (cell, group) pair of the tile, a
// lane per pair, against the weighted direction cosines, which are scalar loads
// (the direction index is the same across the wave); waves 1..3 hand their sums
// to wave 0 through LDS, which adds them in wave order and writes the FP64
// quadrant partial moments.
//
// Faces between tiles go through exchange buffers as one record per (tile,
// lane) -- 2 TY values across a V boundary, 2 TX across an H boundary -- moved
// in 16-byte chunks laid out chunk-major, so every chunk instruction of a wave
// covers 64 consecutive 16-byte records.  The accesses carry the device-scope
// cache policy (sc1) of agent-scope atomics, so they stay coherent across the
// XCDs' L2 caches.  The producer waits for its stores to complete
// (s_waitcnt vmcnt(0)) before the barrier after which one lane publishes the
// completion flag, and the consumer reads the faces only after it has seen the
// flag; no L2 writeback or invalidate is needed for the rest of the data,
// which no other workgroup of the launch reads.
//
// The corner solve, the reflection to the first quadrant and the intensity
// layout [cell][quadrant][group][direction][corner] are those of sweep.hip.
#include "trt_internal.h"

#include <algorithm>

namespace trt {

namespace {

constexpr int PK_THREADS = 256;

struct PkProb {
    int nx, ny, ng, K, gblk, NTX, NTY;
    int kw, npair, xsw;    // moment reduction: directions per wave, (cell, group) pairs per tile, doubles of
                           // the cross-wave sums ahead of the staging area in LDS (0: they reuse it)
    int seg, nseg;         // column segments: tiles per segment, segments per tile column
    double inv_cdt, eps;
    const double *dx, *dy, *ox, *oy, *w;
    const void *sc;        // [cell][g] (KapE + 1/(c dt), RT_Src) in the sweep precision
    const double *bcin;    // [4][ng]
    const void *Iold;
    void *Inew;
    void *XV, *XH;
    double *Part, *PVb, *PHb;  // quadrant partial moments (PartLayout) and their incoming boundary faces
    int PP, nblk_g;            // the layout's run length and group blocks
    int NLr;                   // lanes of a face-exchange or boundary slice (rec_lanes)
    const double *ctab;    // [m][6]: w, w ox^2, w oy^2, w ox oy, w ox, w oy, over the stored intensities' factor
    int tx, ty;            // tile shape
    int refl;              // reflective sides: bit 0 left, 1 bottom, 2 right, 3 top
    int skip;              // batched sweeps: this problem is not swept this time (its items only advance the queue)
    void *BV, *BH;         // their boundary slices: [left, right][q][NTY][NLr][2 TY], [bottom, top][q][NTX][NLr][2 TX]
};

struct PkArgs {
    const PkProb *probs;
    const int4 *items;     // {problem, own flag, txl | tyl << 16, q | gb << 8}
    int nitems;
    int *queue;            // [0] next item, [1] epoch, [2] error word
    int *flags;
    int spin_limit;
};

// Pointers of the problem descriptor are generic once the descriptor has been
// read from memory; accesses through them would be flat instructions, which the
// compiler orders against LDS traffic with full waits.  The tile engine casts
// them to the global address space at the point of use.
#define GLOBAL_AS __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ GLOBAL_AS T *gptr(T *p) { return (GLOBAL_AS T *)p; }

typedef double d2v __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));
template <typename T> struct Vec2;
template <> struct Vec2<double> { typedef d2v type; };
template <> struct Vec2<float> { typedef f2v type; };

template <typename T> struct V4;
template <> struct V4<double> {
    double v[4];
    __device__ static V4 load(GLOBAL_AS const double *p) {
        const d2v a = *reinterpret_cast<GLOBAL_AS const d2v *>(p);
        const d2v b = *reinterpret_cast<GLOBAL_AS const d2v *>(p + 2);
        V4 r; r.v[0] = a.x; r.v[1] = a.y; r.v[2] = b.x; r.v[3] = b.y; return r;
    }
    __device__ void store(GLOBAL_AS double *p) const {
        d2v a, b;
        a.x = v[0]; a.y = v[1]; b.x = v[2]; b.y = v[3];
        *reinterpret_cast<GLOBAL_AS d2v *>(p) = a;
        *reinterpret_cast<GLOBAL_AS d2v *>(p + 2) = b;
    }
};
template <> struct V4<float> {
    float v[4];
    __device__ static V4 load(GLOBAL_AS const float *p) {
        const f4v a = *reinterpret_cast<GLOBAL_AS const f4v *>(p);
        V4 r; r.v[0] = a.x; r.v[1] = a.y; r.v[2] = a.z; r.v[3] = a.w; return r;
    }
    __device__ void store(GLOBAL_AS float *p) const {
        f4v a;
        a.x = v[0]; a.y = v[1]; a.z = v[2]; a.w = v[3];
        *reinterpret_cast<GLOBAL_AS f4v *>(p) = a;
    }
};

__device__ __forceinline__ void store4(GLOBAL_AS double *o, double a, double b, double c, double d)
{
    d2v x, y;
    x.x = a; x.y = b; y.x = c; y.y = d;
    *reinterpret_cast<GLOBAL_AS d2v *>(o) = x;
    *reinterpret_cast<GLOBAL_AS d2v *>(o + 2) = y;
}

// The same pointer in every lane (a value read from LDS the compiler cannot prove
// uniform): keep it scalar, so buffer descriptors built from it stay in SGPRs.
template <typename T>
__device__ __forceinline__ T *uniform_ptr(const T *p)
{
    const unsigned long long a = reinterpret_cast<unsigned long long>(p);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
    return reinterpret_cast<T *>(((unsigned long long)hi << 32) | lo);
}

// Face exchange.  A slice holds the records of one tile boundary line:
// [tile][16-byte chunk][lane]; a record is N values of one lane.  Buffer accesses
// with the SC1 cache-policy bit, the device-scope coherence of agent-scope atomics.
typedef unsigned int u4v __attribute__((ext_vector_type(4)));
constexpr int kBufDword3 = 0x00020000;   // buffer descriptor word 3 of gfx9 (raw, 32-bit elements)
constexpr int kSc1 = 16;                 // cache-policy bit SC1

template <typename TI, int N>
__device__ __forceinline__ void xload_rec(const TI *slice, unsigned bytes, unsigned tile, unsigned NL, unsigned lane,
                                          TI (&v)[N])
{
    static_assert((N * sizeof(TI)) % 16 == 0, "face records are whole 16-byte chunks");
    constexpr int NC = (int)(N * sizeof(TI) / 16);
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
        uniform_ptr(slice), 0, (int)__builtin_amdgcn_readfirstlane(bytes), kBufDword3);
#pragma unroll
    for (int ch = 0; ch < NC; ++ch) {
        const unsigned off = ((tile * NC + ch) * NL + lane) * 16u;
        const u4v w = __builtin_bit_cast(u4v, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, kSc1));
        __builtin_memcpy(reinterpret_cast<char *>(v) + 16 * ch, &w, 16);
    }
}

template <typename TI, int N>
__device__ __forceinline__ void xstore_rec(TI *slice, unsigned bytes, unsigned tile, unsigned NL, unsigned lane,
                                           const TI (&v)[N])
{
    constexpr int NC = (int)(N * sizeof(TI) / 16);
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
        uniform_ptr(slice), 0, (int)__builtin_amdgcn_readfirstlane(bytes), kBufDword3);
#pragma unroll
    for (int ch = 0; ch < NC; ++ch) {
        u4v w;
        __builtin_memcpy(&w, reinterpret_cast<const char *>(v) + 16 * ch, 16);
        const unsigned off = ((tile * NC + ch) * NL + lane) * 16u;
        __builtin_amdgcn_raw_buffer_store_b128(w, r, (int)off, 0, kSc1);
    }
}

// 1/x: hardware reciprocal, refined by Newton-Raphson steps (r <- r + r (1 - x r)),
// each of which squares the relative error: one step for FP32, two for FP64.
// The refinement removes the bias of the hardware estimate, which would
// otherwise accumulate along the sweep's characteristics.
__device__ __forceinline__ float recip(float x)
{
    const float r = __builtin_amdgcn_rcpf(x);
    return fmaf(r, fmaf(-x, r, 1.0f), r);
}
__device__ __forceinline__ double recip(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// s_waitcnt vmcnt(0): every global load and store of this wave has completed.  The
// compiler barriers keep every store of the tile (the face records are buffer
// stores the compiler does not order against the wait by itself) ahead of it.
__device__ __forceinline__ void wait_vm_stores()
{
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0x0F70);
    asm volatile("" ::: "memory");
}

// Bounded wait for an upwind tile.  A wait that runs out sets the error word, and
// every later wait of the launch sees it and returns at once, so a broken
// dependency ends the launch within one spin limit instead of hanging it.
__device__ __forceinline__ void wait_flag(const int *f, int epoch, int limit, int *err)
{
    int n = 0;
    while (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
        __builtin_amdgcn_s_sleep(2);
        ++n;
        if ((n & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
        if (n > limit) {
            atomicOr(err, 1);
            break;
        }
    }
}

// Sweep one TX x TY tile of quadrant q (signs px, py) for group block gb; stage
// the corner averages (cell, downwind V face, downwind H face) in st.  The loads
// that do not depend on other tiles (old intensities, coefficients, geometry)
// are issued before sync(), which waits for the upwind tiles; the faces those
// tiles wrote are read after it.
// hc carries the H faces between consecutive tiles of one column swept by the same
// workgroup: with hin_regs the tile's incoming H faces come from hc instead of the
// exchange buffer, and with hout_xh false its outgoing ones stay in hc only.  The
// carried values are rounded to TI exactly as the exchange buffer rounds them.
// publish() runs once the outgoing faces are stored (it makes them visible to the
// downwind tiles); the tile's new corner intensities are kept in registers until then
// and stored after it, so their stores are not on the path to the downwind tiles.
template <typename TI, typename TC, int TX, int TY, bool px, bool py, typename Sync, typename Publish>
__device__ __forceinline__ void tile_body(const PkProb &P, TI *st, int q, int gb, int txl, int tyl, Sync sync,
                                          Publish publish, TI (&hc)[TX][2], bool hin_regs, bool hout_xh)
{
    const int tid = threadIdx.x;
    const int K = P.K, ng = P.ng, nx = P.nx, ny = P.ny;
    const int NL = ng * K;
    const int gl = tid / K, k = tid - gl * K;
    const int g = gb * P.gblk + gl;
    const bool act = (gl < P.gblk) && (g < ng);
    const int L = act ? g * K + k : 0;
    const int m = q * K + k;
    const double axn = fabs(gptr(P.ox)[m]), ayn = fabs(gptr(P.oy)[m]);
    constexpr int pmask = (px ? 0 : 1) | (py ? 0 : 2);
    constexpr int perm[4] = {pmask == 0 ? 0 : pmask == 1 ? 1 : pmask == 2 ? 3 : 2,
                             pmask == 0 ? 1 : pmask == 1 ? 0 : pmask == 2 ? 2 : 3,
                             pmask == 0 ? 2 : pmask == 1 ? 3 : pmask == 2 ? 1 : 0,
                             pmask == 0 ? 3 : pmask == 1 ? 2 : pmask == 2 ? 0 : 1};
    GLOBAL_AS const TI *Iold = gptr(static_cast<const TI *>(P.Iold));
    GLOBAL_AS TI *Inew = gptr(static_cast<TI *>(P.Inew));
    TI *XV = static_cast<TI *>(P.XV);
    TI *XH = static_cast<TI *>(P.XH);
    typedef typename Vec2<TI>::type TI2;
    GLOBAL_AS const TI2 *scv = reinterpret_cast<GLOBAL_AS const TI2 *>(gptr(P.sc));
    const TC inv_cdt = P.inv_cdt;
    const int i0 = txl * TX, iw = min(TX, nx - i0);
    const int j0 = tyl * TY, jh = min(TY, ny - j0);
    const int gg = act ? g : 0;

    // ---- every load of the tile, issued up front (the values are used after all
    // loads have been issued, so the tile costs one memory round trip) ----
    double dyv[TY], dxv[TX];
#pragma unroll
    for (int r = 0; r < TY; ++r) {
        const int j = py ? j0 + r : ny - 1 - (j0 + r);
        dyv[r] = r < jh ? gptr(P.dy)[j] : 0.0;
    }
#pragma unroll
    for (int ii = 0; ii < TX; ++ii) {
        const int i = px ? i0 + ii : nx - 1 - (i0 + ii);
        dxv[ii] = ii < iw ? gptr(P.dx)[i] : 0.0;
    }
    V4<TI> io[TY][TX];
    TI kt[TY][TX], sr[TY][TX];
#pragma unroll
    for (int r = 0; r < TY; ++r) {
        const int j = py ? j0 + r : ny - 1 - (j0 + r);
#pragma unroll
        for (int ii = 0; ii < TX; ++ii) {
            if (r < jh && ii < iw && act) {
                const int i = px ? i0 + ii : nx - 1 - (i0 + ii);
                const size_t c = (size_t)i + (size_t)nx * j;
                io[r][ii] = V4<TI>::load(Iold + ((c * 4 + q) * NL + L) * 4);
                const TI2 s = scv[c * ng + gg];
                kt[r][ii] = s.x;
                sr[r][ii] = s.y;
            } else {
                io[r][ii].v[0] = io[r][ii].v[1] = io[r][ii].v[2] = io[r][ii].v[3] = (TI)0;
                kt[r][ii] = (TI)1;
                sr[r][ii] = (TI)0;
            }
        }
    }
    const TI bcx = (TI)(gptr(P.bcin)[(px ? 0 : 2) * ng + gg] * iscale<TI>());
    const TI bcy = (TI)(gptr(P.bcin)[(py ? 1 : 3) * ng + gg] * iscale<TI>());

    sync();
    asm volatile("" ::: "memory");   // the face loads stay behind the upwind waits

    // incoming faces: the records the left and lower tiles wrote (rows and columns
    // beyond a partial tile's edge are carried but unused)
    constexpr int NV = 2 * TY, NH = 2 * TX;
    const int NR = P.NLr;
    const unsigned vslice = (unsigned)((size_t)P.NTY * NR * NV * sizeof(TI));
    const unsigned hslice = (unsigned)((size_t)P.NTX * NR * NH * sizeof(TI));
    TI vbr[TY], vtr[TY], hl[TX], hr[TX];
    {
        TI pv[NV];
#pragma unroll
        for (int t = 0; t < NV; ++t) pv[t] = bcx;
        if (txl > 0 && act)
            xload_rec<TI, NV>(XV + (size_t)(q * (P.NTX - 1) + txl - 1) * P.NTY * NR * NV, vslice, (unsigned)tyl,
                              (unsigned)NR, (unsigned)L, pv);
        else if (act && (P.refl & (px ? 1 : 4)))   // reflective left (right) side: the inflow slice
            xload_rec<TI, NV>(static_cast<TI *>(P.BV) + (size_t)((px ? 0 : 1) * 4 + q) * P.NTY * NR * NV, vslice,
                              (unsigned)tyl, (unsigned)NR, (unsigned)L, pv);
#pragma unroll
        for (int r = 0; r < TY; ++r) { vbr[r] = pv[2 * r]; vtr[r] = pv[2 * r + 1]; }
        TI ph[NH];
#pragma unroll
        for (int t = 0; t < NH; ++t) ph[t] = bcy;
        if (tyl > 0 && act && !hin_regs)
            xload_rec<TI, NH>(XH + (size_t)(q * (P.NTY - 1) + tyl - 1) * P.NTX * NR * NH, hslice, (unsigned)txl,
                              (unsigned)NR, (unsigned)L, ph);
        else if (tyl == 0 && act && (P.refl & (py ? 2 : 8)))   // reflective bottom (top) side
            xload_rec<TI, NH>(static_cast<TI *>(P.BH) + (size_t)((py ? 0 : 1) * 4 + q) * P.NTX * NR * NH, hslice,
                              (unsigned)txl, (unsigned)NR, (unsigned)L, ph);
#pragma unroll
        for (int ii = 0; ii < TX; ++ii) {
            hl[ii] = ph[2 * ii];
            hr[ii] = ph[2 * ii + 1];
            if (tyl > 0 && hin_regs) { hl[ii] = hc[ii][0]; hr[ii] = hc[ii][1]; }
        }
    }

    // ---- geometry of the tile's rows and columns for this direction ----
    // aa = |Omega_x| dy / 4, bb = |Omega_y| dx / 4, A/4 = dx dy / 4 (SCB_BGEN, :332-371)
    TC aar[TY], dyr[TY], bbc[TX], dx4c[TX];
#pragma unroll
    for (int r = 0; r < TY; ++r) {
        dyr[r] = (TC)dyv[r];
        aar[r] = (TC)(axn * dyv[r] * 0.25);
    }
#pragma unroll
    for (int ii = 0; ii < TX; ++ii) {
        dx4c[ii] = (TC)(dxv[ii] * 0.25);
        bbc[ii] = (TC)(ayn * dxv[ii] * 0.25);
    }
    TC hlc[TX], hrc[TX];
#pragma unroll
    for (int ii = 0; ii < TX; ++ii) { hlc[ii] = (TC)hl[ii]; hrc[ii] = (TC)hr[ii]; }

    // ---- cells in upwind order; the corner solve runs in the compute type TC,
    // values are rounded to TI where they are stored ----
#pragma unroll
    for (int r = 0; r < TY; ++r) {
        if (r >= jh) continue;
        const int j = py ? j0 + r : ny - 1 - (j0 + r);
        TC vb = (TC)vbr[r], vt = (TC)vtr[r];
        const TC aa = aar[r];
#pragma unroll
        for (int ii = 0; ii < TX; ++ii) {
            if (ii >= iw) continue;
            const TC bb = bbc[ii];
            const TC A4 = dx4c[ii] * dyr[r];
            const TC gam = aa + bb + (TC)kt[r][ii] * A4;
            const TC src = (TC)sr[r][ii];
            // S = RT_Src + I_old / (c dt)  (TRANSPORT_SOLVES.f08:48); Q = B x (:270-294)
            const TC s0 = src + (TC)io[r][ii].v[perm[0]] * inv_cdt;
            const TC s1 = src + (TC)io[r][ii].v[perm[1]] * inv_cdt;
            const TC s2 = src + (TC)io[r][ii].v[perm[2]] * inv_cdt;
            const TC s3 = src + (TC)io[r][ii].v[perm[3]] * inv_cdt;
            const TC Q0 = A4 * s0 + (TC)2 * aa * vb + (TC)2 * bb * hlc[ii];
            const TC Q1 = A4 * s1 + (TC)2 * bb * hrc[ii];
            const TC Q2 = A4 * s2;
            const TC Q3 = A4 * s3 + (TC)2 * aa * vt;
            const TC g2 = gam * gam;
            const TC al = g2 + aa * aa + bb * bb, be = (TC)2 * aa * bb;
            const TC y0 = al * Q0 + be * Q2, y1 = al * Q1 - be * Q3;
            const TC y2 = al * Q2 + be * Q0, y3 = al * Q3 - be * Q1;
            const TC dm = aa - bb, dp = aa + bb;
            const TC rD = recip((g2 + dm * dm) * (g2 + dp * dp));
            const TC x0 = (gam * y0 - aa * y1 - bb * y3) * rD;
            const TC x1 = (gam * y1 + aa * y0 - bb * y2) * rD;
            const TC x2 = (gam * y2 + bb * y1 + aa * y3) * rD;
            const TC x3 = (gam * y3 + bb * y0 - aa * y2) * rD;
            // the old corners of the cell are no longer needed: its new ones take their registers
            io[r][ii].v[perm[0]] = (TI)x0; io[r][ii].v[perm[1]] = (TI)x1;
            io[r][ii].v[perm[2]] = (TI)x2; io[r][ii].v[perm[3]] = (TI)x3;
            // outgoing corner edges (:86-89)
            vb = x1; vt = x2; hlc[ii] = x3; hrc[ii] = x2;
            // corner averages for COLLAPSE_INTENSITIES (:447-470)
            const int cidx = r * TX + ii;
            st[(0 * TX * TY + cidx) * PK_THREADS + tid] = (TI)((x0 + x1 + x2 + x3) * (TC)0.25);
            st[(1 * TX * TY + cidx) * PK_THREADS + tid] = (TI)((x1 + x2) * (TC)0.5);
            st[(2 * TX * TY + cidx) * PK_THREADS + tid] = (TI)((x3 + x2) * (TC)0.5);
        }
        vbr[r] = (TI)vb;                 // outgoing V faces of the row
        vtr[r] = (TI)vt;
    }
    // outgoing faces: to the next tile, or on a reflective side of the domain to its
    // outflow slice (RT_BC_UPDATE copies it to the mirror quadrant's inflow after the sweep)
    const bool v_next = txl < P.NTX - 1, v_bnd = !v_next && (P.refl & (px ? 4 : 1));
    if ((v_next || v_bnd) && act) {
        TI pv[NV];
#pragma unroll
        for (int r = 0; r < TY; ++r) { pv[2 * r] = vbr[r]; pv[2 * r + 1] = vtr[r]; }
        TI *dst = v_next ? XV + (size_t)(q * (P.NTX - 1) + txl) * P.NTY * NR * NV
                     : static_cast<TI *>(P.BV) + (size_t)((px ? 1 : 0) * 4 + q) * P.NTY * NR * NV;
        xstore_rec<TI, NV>(dst, vslice, (unsigned)tyl, (unsigned)NR, (unsigned)L, pv);
    }
    const bool h_next = tyl < P.NTY - 1 && hout_xh, h_bnd = tyl == P.NTY - 1 && (P.refl & (py ? 8 : 2));
    if ((h_next || h_bnd) && act) {
        TI ph[NH];
#pragma unroll
        for (int ii = 0; ii < TX; ++ii) { ph[2 * ii] = (TI)hlc[ii]; ph[2 * ii + 1] = (TI)hrc[ii]; }
        TI *dst = h_next ? XH + (size_t)(q * (P.NTY - 1) + tyl) * P.NTX * NR * NH
                     : static_cast<TI *>(P.BH) + (size_t)((py ? 1 : 0) * 4 + q) * P.NTX * NR * NH;
        xstore_rec<TI, NH>(dst, hslice, (unsigned)txl, (unsigned)NR, (unsigned)L, ph);
    }
#pragma unroll
    for (int ii = 0; ii < TX; ++ii) {
        hc[ii][0] = (TI)hlc[ii];
        hc[ii][1] = (TI)hrc[ii];
    }
    publish();
    // the tile's I_crn (TRANSPORT_SOLVES.f08:95-98)
#pragma unroll
    for (int r = 0; r < TY; ++r) {
        const int j = py ? j0 + r : ny - 1 - (j0 + r);
#pragma unroll
        for (int ii = 0; ii < TX; ++ii) {
            if (r < jh && ii < iw && act) {
                const int i = px ? i0 + ii : nx - 1 - (i0 + ii);
                io[r][ii].store(Inew + ((((size_t)i + (size_t)nx * j) * 4 + q) * NL + L) * 4);
            }
        }
    }
}

typedef const __attribute__((address_space(4))) double CDBL;

// Sum the staged corner averages of a tile over the directions of each group and
// write the quadrant partial moments of its cells and downwind faces.  Wave w takes
// directions [w kw, (w+1) kw), a lane per (cell, group) pair (64 pairs per pass); the
// direction index is uniform across the wave, so the weights are scalar loads from
// the problem's constant table.  Waves 1..3 leave their 13 sums in xs; wave 0 adds
// them in wave order (a fixed order: results do not depend on timing) and writes the
// partials.  With P.xsw == 0 the sums of at most two passes are formed before xs
// overwrites the staging area; otherwise xs has its own LDS area, one pass at a time.
constexpr int MAX_OVERLAP_PASSES = 2;

template <typename TS, int TX, int TY, bool px, bool py>
__device__ __forceinline__ void reduce_tile(const PkProb &P, const TS *st, double *xs, int q, int gb, int txl, int tyl)
{
    const int tid = threadIdx.x, ln = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int K = P.K, ng = P.ng, nx = P.nx, ny = P.ny, gblk = P.gblk;
    const int k0 = wv * P.kw, k1 = min(K, k0 + P.kw);
    const int i0 = txl * TX, iw = min(TX, nx - i0);
    const int j0 = tyl * TY, jh = min(TY, ny - j0);
    const int g0 = gb * gblk;
    const double eps = P.eps * iscale<TS>();   // (the weights carry 1 / iscale)
    CDBL *ct = (CDBL *)(uniform_ptr(P.ctab)) + (size_t)q * K * 6;
    // pair p of the tile: (cell cidx = (r, ii), group g); ok if it lies inside the grid
    auto pair_of = [&](int p, int &cidx, int &gl, int &r, int &ii, int &g) {
        cidx = p / gblk; gl = p - cidx * gblk;
        r = cidx / TX; ii = cidx - r * TX;
        g = g0 + gl;
        return p < P.npair && ii < iw && r < jh && g < ng;
    };
    auto accumulate = [&](int p, double (&m)[NMOM]) {
#pragma unroll
        for (int v = 0; v < NMOM; ++v) m[v] = 0.0;
        int cidx, gl, r, ii, g;
        if (!pair_of(p, cidx, gl, r, ii, g)) return;
        const TS *sa = st + (0 * TX * TY + cidx) * PK_THREADS + gl * K;
        const TS *sv = st + (1 * TX * TY + cidx) * PK_THREADS + gl * K;
        const TS *sh = st + (2 * TX * TY + cidx) * PK_THREADS + gl * K;
        for (int kk = k0; kk < k1; ++kk) {
            const double ia = (double)sa[kk] + eps, iv = (double)sv[kk] + eps, ih = (double)sh[kk] + eps;
            CDBL *cc = ct + kk * 6;   // w, w ox^2, w oy^2, w ox oy, w ox, w oy
            m[0] += cc[1] * ia; m[1] += cc[2] * ia; m[2] += cc[3] * ia; m[3] += cc[0] * ia;
            m[4] += cc[1] * iv; m[5] += cc[3] * iv; m[6] += cc[4] * iv; m[7] += cc[0] * iv;
            m[8] += cc[2] * ih; m[9] += cc[3] * ih; m[10] += cc[4] * ih; m[11] += cc[5] * ih;
            m[12] += cc[0] * ih;
        }
    };
    auto finish = [&](int p, const double *xp, double (&m)[NMOM]) {   // wave 0: add waves 1..3, store
        int cidx, gl, r, ii, g;
        if (!pair_of(p, cidx, gl, r, ii, g)) return;
#pragma unroll
        for (int w = 1; w < 4; ++w)
#pragma unroll
            for (int v = 0; v < NMOM; ++v) m[v] += xp[((w - 1) * NMOM + v) * 64 + ln];
        // the item's runs: value v of pair p at v PP + p, so lane ln of the wave writes the
        // value next to lane ln - 1's
        GLOBAL_AS double *o = gptr(P.Part) + (size_t)P.PP * NMOM *
                                                 ((((size_t)q * P.nblk_g + gb) * P.NTY + tyl) * P.NTX + txl) + p;
#pragma unroll
        for (int v = 0; v < NMOM; ++v) o[(size_t)v * P.PP] = m[v];
    };
    constexpr int XSP = 3 * NMOM * 64;   // doubles of one pass in xs
    if (P.xsw == 0) {
        const int npass = (P.npair + 63) / 64;   // at most MAX_OVERLAP_PASSES here
        double m[MAX_OVERLAP_PASSES][NMOM];
#pragma unroll
        for (int t = 0; t < MAX_OVERLAP_PASSES; ++t) accumulate(t < npass ? t * 64 + ln : P.npair, m[t]);
        __syncthreads();                 // every wave has read the staged values xs overwrites
        if (wv > 0) {
#pragma unroll
            for (int t = 0; t < MAX_OVERLAP_PASSES; ++t)
                if (t < npass)
#pragma unroll
                    for (int v = 0; v < NMOM; ++v) xs[t * XSP + ((wv - 1) * NMOM + v) * 64 + ln] = m[t][v];
        }
        __syncthreads();
        if (wv == 0) {
#pragma unroll
            for (int t = 0; t < MAX_OVERLAP_PASSES; ++t)
                if (t < npass) finish(t * 64 + ln, xs + t * XSP, m[t]);
        }
    } else {
        for (int pb = 0; pb < P.npair; pb += 64) {
            double m[NMOM];
            accumulate(pb + ln, m);
            if (wv > 0) {
#pragma unroll
                for (int v = 0; v < NMOM; ++v) xs[((wv - 1) * NMOM + v) * 64 + ln] = m[v];
            }
            __syncthreads();
            if (wv == 0) finish(pb + ln, xs, m);
            if (pb + 64 < P.npair) __syncthreads();   // xs is written again by the next pass
        }
    }
}

template <typename TI, typename TC, int TX, int TY, typename Sync, typename Publish>
__device__ __forceinline__ void tile_sweep(const PkProb &P, TI *st, int q, int gb, int txl, int tyl, Sync sync,
                                           Publish publish, TI (&hc)[TX][2], bool hin_regs, bool hout_xh)
{
    switch (q) {
    case 0: tile_body<TI, TC, TX, TY, true, true>(P, st, q, gb, txl, tyl, sync, publish, hc, hin_regs, hout_xh); break;
    case 1: tile_body<TI, TC, TX, TY, false, true>(P, st, q, gb, txl, tyl, sync, publish, hc, hin_regs, hout_xh); break;
    case 2: tile_body<TI, TC, TX, TY, false, false>(P, st, q, gb, txl, tyl, sync, publish, hc, hin_regs, hout_xh);
        break;
    default: tile_body<TI, TC, TX, TY, true, false>(P, st, q, gb, txl, tyl, sync, publish, hc, hin_regs, hout_xh);
        break;
    }
}

template <typename TI, int TX, int TY>
__device__ __forceinline__ void tile_reduce(const PkProb &P, const TI *st, double *xs, int q, int gb, int txl, int tyl)
{
    switch (q) {
    case 0: reduce_tile<TI, TX, TY, true, true>(P, st, xs, q, gb, txl, tyl); break;
    case 1: reduce_tile<TI, TX, TY, false, true>(P, st, xs, q, gb, txl, tyl); break;
    case 2: reduce_tile<TI, TX, TY, false, false>(P, st, xs, q, gb, txl, tyl); break;
    default: reduce_tile<TI, TX, TY, true, false>(P, st, xs, q, gb, txl, tyl); break;
    }
}

// The problem descriptor of the current item is copied into LDS once per item:
// read from global memory at every use, each pointer field would cost a vector
// load and a wait in front of the loads it feeds.
constexpr int PKPROB_WORDS = (int)(sizeof(PkProb) / 8);
static_assert(sizeof(PkProb) % 8 == 0, "PkProb is copied as 8-byte words");

// Waves per SIMD the persistent tile kernel is compiled for: the FP32 4x4 tile needs a
// few registers above the budget of 3 (168 VGPRs) and is held to it without spilling;
// the others keep the compiler's choice (the column-segment kernel would spill).
template <typename TI, typename TC, int TX, int TY>
struct PkOccupancy {
    static constexpr int waves = (sizeof(TI) == 4 && sizeof(TC) == 4 && TX == 4 && TY == 4) ? 3 : 1;
};

template <typename TI, typename TC, int TX, int TY>
__global__ __launch_bounds__(PK_THREADS, (PkOccupancy<TI, TC, TX, TY>::waves)) void sweep_pk(PkArgs a)
{
    extern __shared__ double dyn[];
    __shared__ int s_next;
    __shared__ int4 s_item;                          // work-list entry of s_next
    __shared__ int s_pi;                             // problem whose descriptor is in s_prob
    __shared__ unsigned long long s_prob[PKPROB_WORDS];
    const int tid = threadIdx.x;
    const int epoch = a.queue[1];
    int *err = a.queue + 2;
    if (tid == 0) {
        const int c0 = atomicAdd(a.queue, 1);
        s_next = c0;
        if (c0 < a.nitems) s_item = a.items[c0];
        s_pi = -1;
    }
    __syncthreads();
    int cur = __builtin_amdgcn_readfirstlane(s_next);
    while (cur < a.nitems) {
        // the next item and its work-list entry are fetched while this one runs
        int nxt = 0;
        int4 nit = make_int4(0, 0, 0, 0);
        if (tid == 0) {
            nxt = atomicAdd(a.queue, 1);
            if (nxt < a.nitems) nit = a.items[nxt];
        }
        const int pi = __builtin_amdgcn_readfirstlane(s_item.x);
        const int fl = __builtin_amdgcn_readfirstlane(s_item.y);
        const int tpos = __builtin_amdgcn_readfirstlane(s_item.z);
        const int qg = __builtin_amdgcn_readfirstlane(s_item.w);
        const int txl = tpos & 0xffff, tyl = tpos >> 16, q = qg & 0xff, gb = qg >> 8;
        if (pi != __builtin_amdgcn_readfirstlane(s_pi)) {
            const PkProb &Pg = a.probs[pi];
            if (tid < PKPROB_WORDS) s_prob[tid] = reinterpret_cast<const unsigned long long *>(&Pg)[tid];
            __syncthreads();
            if (tid == 0) s_pi = pi;
        }
        const PkProb &P = *reinterpret_cast<const PkProb *>(s_prob);
        if (P.skip) {                     // a problem left out of this sweep: no wait, no work
            __syncthreads();              // every wave has read s_item
            if (tid == 0) {
                s_next = nxt;
                s_item = nit;
            }
            __syncthreads();
            cur = __builtin_amdgcn_readfirstlane(s_next);
            continue;
        }
        double *xs = dyn;                                           // cross-wave sums
        TI *st = reinterpret_cast<TI *>(dyn + P.xsw);             // staged corner averages
        if (P.xsw == 0) xs = reinterpret_cast<double *>(st);
        TI hc[TX][2];
        tile_sweep<TI, TC, TX, TY>(P, st, q, gb, txl, tyl, [&]() {
            if (tid == 0) {
                if (txl > 0) wait_flag(a.flags + fl - 1, epoch, a.spin_limit, err);
                if (tyl > 0) wait_flag(a.flags + fl - P.NTX, epoch, a.spin_limit, err);
            }
            __syncthreads();
        }, [&]() {
            wait_vm_stores();             // this wave's face stores have completed
            __syncthreads();
            if (tid == 0) {
                __hip_atomic_store(a.flags + fl, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                s_next = nxt;
                s_item = nit;
            }
        }, hc, false, true);
        tile_reduce<TI, TX, TY>(P, st, xs, q, gb, txl, tyl);
        __syncthreads();
        cur = __builtin_amdgcn_readfirstlane(s_next);
    }
}

// One launch per tile diagonal d: the tiles of d of every quadrant and group block.
template <typename TI, typename TC, int TX, int TY>
__global__ __launch_bounds__(PK_THREADS) void sweep_diag(const PkProb P, int d, int tx_lo, int nt_d, int nblk_g)
{
    extern __shared__ double dyn[];
    int bid = blockIdx.x;
    const int gb = bid % nblk_g;
    bid /= nblk_g;
    const int t = bid % nt_d;
    const int q = bid / nt_d;
    const int txl = tx_lo + t, tyl = d - txl;
    double *xs = dyn;
    TI *st = reinterpret_cast<TI *>(dyn + P.xsw);
    if (P.xsw == 0) xs = reinterpret_cast<double *>(st);
    TI hc[TX][2];
    tile_sweep<TI, TC, TX, TY>(P, st, q, gb, txl, tyl, [&]() { __syncthreads(); }, [&]() {}, hc, false, true);
    __syncthreads();
    tile_reduce<TI, TX, TY>(P, st, xs, q, gb, txl, tyl);
}

// Column-segment form: a work item is (quadrant, group block, tile column, segment
// of seg consecutive tiles of that column).  The workgroup sweeps the segment's
// tiles bottom to top and keeps the H faces between them in registers, so only
// the segment's last tile writes the H exchange buffer and only its first tile
// waits for the tile below; every tile still waits for its left neighbour and
// publishes its own flag.  Items are listed by segment diagonal (tile column +
// segment index), a topological order of the segments.
template <typename TI, typename TC, int TX, int TY>
__global__ __launch_bounds__(PK_THREADS) void sweep_seg(PkArgs a)
{
    extern __shared__ double dyn[];
    __shared__ int s_next;
    __shared__ int4 s_item;
    __shared__ int s_pi;
    __shared__ unsigned long long s_prob[PKPROB_WORDS];
    const int tid = threadIdx.x;
    const int epoch = a.queue[1];
    int *err = a.queue + 2;
    if (tid == 0) {
        const int c0 = atomicAdd(a.queue, 1);
        s_next = c0;
        if (c0 < a.nitems) s_item = a.items[c0];
        s_pi = -1;
    }
    __syncthreads();
    int cur = __builtin_amdgcn_readfirstlane(s_next);
    while (cur < a.nitems) {
        int nxt = 0;
        int4 nit = make_int4(0, 0, 0, 0);
        if (tid == 0) {
            nxt = atomicAdd(a.queue, 1);
            if (nxt < a.nitems) nit = a.items[nxt];
        }
        const int pi = __builtin_amdgcn_readfirstlane(s_item.x);
        const int fl0 = __builtin_amdgcn_readfirstlane(s_item.y);
        const int tpos = __builtin_amdgcn_readfirstlane(s_item.z);
        const int qg = __builtin_amdgcn_readfirstlane(s_item.w);
        const int txl = tpos & 0xffff, sg = tpos >> 16, q = qg & 0xff, gb = qg >> 8;
        if (pi != __builtin_amdgcn_readfirstlane(s_pi)) {
            const PkProb &Pg = a.probs[pi];
            if (tid < PKPROB_WORDS) s_prob[tid] = reinterpret_cast<const unsigned long long *>(&Pg)[tid];
            __syncthreads();
            if (tid == 0) s_pi = pi;
        }
        const PkProb &P = *reinterpret_cast<const PkProb *>(s_prob);
        double *xs = dyn;
        TI *st = reinterpret_cast<TI *>(dyn + P.xsw);
        if (P.xsw == 0) xs = reinterpret_cast<double *>(st);
        const int t0 = sg * P.seg, t1 = P.skip ? t0 : min(t0 + P.seg, P.NTY);   // a skipped problem: no tiles
        TI hc[TX][2];
        for (int tyl = t0; tyl < t1; ++tyl) {
            const int fl = fl0 + (tyl - t0) * P.NTX;
            tile_sweep<TI, TC, TX, TY>(P, st, q, gb, txl, tyl, [&]() {
                if (tid == 0) {
                    if (txl > 0) wait_flag(a.flags + fl - 1, epoch, a.spin_limit, err);
                    if (tyl == t0 && tyl > 0) wait_flag(a.flags + fl - P.NTX, epoch, a.spin_limit, err);
                }
                __syncthreads();
            }, [&]() {
                wait_vm_stores();         // this wave's face stores have completed
                __syncthreads();
                if (tid == 0) __hip_atomic_store(a.flags + fl, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }, hc, tyl > t0, tyl == t1 - 1);
            tile_reduce<TI, TX, TY>(P, st, xs, q, gb, txl, tyl);
            __syncthreads();
        }
        if (P.skip) __syncthreads();      // every wave has read s_item
        if (tid == 0) {
            s_next = nxt;
            s_item = nit;
        }
        __syncthreads();
        cur = __builtin_amdgcn_readfirstlane(s_next);
    }
}

__global__ void pk_reset(int *queue)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        queue[0] = 0;
        queue[1] = queue[1] + 1;
    }
}

// Incoming boundary faces: their moments are those of the boundary intensities
// (RT_BC_UPDATE, MainSrc/UPDATES.f08:12-105): Bg_planck(bcT), the same for every face
// of a side, or on a reflective side the face averages of the inflow slice's corner
// values (I_edgV = (Ic_bottom + Ic_top) / 2 of the mirror direction, TRANSPORT_SOLVES.f08:93).
template <typename TI>
__device__ __forceinline__ double inflow_face(const TI *slice, int tile, int t, int N, int NL, int lane)
{
    const double a = (double)slice[rec_elem<TI>(tile, t, N, NL, lane)];
    const double b = (double)slice[rec_elem<TI>(tile, t + 1, N, NL, lane)];
    return (double)(TI)((a + b) * 0.5) / iscale<TI>();
}

template <typename TI>
__global__ void pk_boundary_moments(const PkProb *probs)
{
    const PkProb &P = probs[blockIdx.y];
    if (P.skip) return;
    const int nx = P.nx, ny = P.ny, ng = P.ng, K = P.K;
    const int per = ny + nx;
    const int tot = 4 * ng * per;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += gridDim.x * blockDim.x) {
        const int q = t / (ng * per);
        const int r = t - q * ng * per;
        const int g = r / per, f = r - g * per;
        const bool px = (q == 0 || q == 3), py = (q == 0 || q == 1);
        if (f < ny) {
            const int j = f;
            const double b0 = (double)(TI)(P.bcin[(px ? 0 : 2) * ng + g] * iscale<TI>()) / iscale<TI>() + P.eps;
            const bool rf = P.refl & (px ? 1 : 4);
            const TI *bs = rf ? static_cast<const TI *>(P.BV) + (size_t)((px ? 0 : 1) * 4 + q) * P.NTY * P.NLr * 2 * P.ty
                              : nullptr;
            const int jf = py ? j : ny - 1 - j, tile = jf / P.ty;
            double s[4] = {0.0, 0.0, 0.0, 0.0};
            for (int k = 0; k < K; ++k) {
                const int m = q * K + k;
                const double b = rf ? inflow_face(bs, tile, 2 * (jf - tile * P.ty), 2 * P.ty, P.NLr, g * K + k) + P.eps
                                    : b0;
                const double wm = P.w[m], ox = P.ox[m], oy = P.oy[m];
                s[0] += wm * ox * ox * b; s[1] += wm * ox * oy * b; s[2] += wm * ox * b; s[3] += wm * b;
            }
            double *o = P.PVb + (((size_t)q * ny + j) * ng + g) * 4;
            o[0] = s[0]; o[1] = s[1]; o[2] = s[2]; o[3] = s[3];
        } else {
            const int i = f - ny;
            const double b0 = (double)(TI)(P.bcin[(py ? 1 : 3) * ng + g] * iscale<TI>()) / iscale<TI>() + P.eps;
            const bool rf = P.refl & (py ? 2 : 8);
            const TI *bs = rf ? static_cast<const TI *>(P.BH) + (size_t)((py ? 0 : 1) * 4 + q) * P.NTX * P.NLr * 2 * P.tx
                              : nullptr;
            const int xf = px ? i : nx - 1 - i, tile = xf / P.tx;
            double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
            for (int k = 0; k < K; ++k) {
                const int m = q * K + k;
                const double b = rf ? inflow_face(bs, tile, 2 * (xf - tile * P.tx), 2 * P.tx, P.NLr, g * K + k) + P.eps
                                    : b0;
                const double wm = P.w[m], ox = P.ox[m], oy = P.oy[m];
                s[0] += wm * oy * oy * b; s[1] += wm * ox * oy * b; s[2] += wm * ox * b; s[3] += wm * oy * b;
                s[4] += wm * b;
            }
            double *o = P.PHb + (((size_t)q * nx + i) * ng + g) * 5;
            o[0] = s[0]; o[1] = s[1]; o[2] = s[2]; o[3] = s[3]; o[4] = s[4];
        }
    }
}

// RT_BC_UPDATE with BC_Type 1 (UPDATES.f08:34-44, 53-63, 72-82, 91-101): on each
// reflective side the incoming quadrants take the last sweep's outflow of the quadrant
// mirrored across the side, direction by direction (the same index within the quadrant),
// face by face and corner by corner.  A quadrant and its mirror across a side share the
// frame along that side (the same tile rows for a V side, the same tile columns for an H
// side, the same corner order in a record), so the update copies whole slices.  In the
// reference's 1-based quadrant numbering: left q1 <- q2, q4 <- q3; bottom q1 <- q4,
// q2 <- q3; right q2 <- q1, q3 <- q4; top q4 <- q1, q3 <- q2.
__device__ __forceinline__ void mirror_pair(int side, int h, int &dst, int &src)
{
    switch (side) {
    case 0: dst = h ? 3 : 0; src = h ? 2 : 1; break;
    case 1: dst = h ? 1 : 0; src = h ? 2 : 3; break;
    case 2: dst = h ? 2 : 1; src = h ? 3 : 0; break;
    default: dst = h ? 2 : 3; src = h ? 1 : 0; break;
    }
}

// blockIdx.y = 2 side + pair; nv, nh: values per V and per H slice
template <typename TI>
__global__ void reflect_kernel(TI *BV, TI *BH, size_t nv, size_t nh, int refl)
{
    const int side = blockIdx.y >> 1, h = blockIdx.y & 1;
    if (!(refl & (1 << side))) return;
    int dst, src;
    mirror_pair(side, h, dst, src);
    const bool vert = side == 0 || side == 2;
    const size_t n = vert ? nv : nh;
    TI *base = vert ? BV + (size_t)(side == 0 ? 0 : 4) * nv : BH + (size_t)(side == 1 ? 0 : 4) * nh;
    const TI *from = base + (size_t)src * n;
    TI *to = base + (size_t)dst * n;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        to[e] = from[e];
}

// Every quadrant's slice of each reflective side (blockIdx.y) at Bg_planck(bcT of the
// side) of the lane's group, as the tile engine rounds the Planckian inflow.
template <typename TI>
__global__ void reflect_reset_kernel(TI *BV, TI *BH, const double *bcin, size_t nv, size_t nh, int NL, int NLr,
                                     int K, int ng, int refl)
{
    const int side = blockIdx.y;
    if (!(refl & (1 << side))) return;
    const bool vert = side == 0 || side == 2;
    const size_t n = 4 * (vert ? nv : nh);
    TI *base = vert ? BV + (size_t)(side == 0 ? 0 : 4) * nv : BH + (size_t)(side == 1 ? 0 : 4) * nh;
    constexpr int E = 16 / (int)sizeof(TI);
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const int lane = (int)((e / E) % NLr);      // lanes from NL on are padding
        base[e] = lane < NL ? (TI)(bcin[side * ng + lane / K] * iscale<TI>()) : (TI)0;
    }
}

// Tile shapes of the tile engine: the staged corner averages of a tile
// (3 TX TY 256 values) fit 48 KiB of LDS and the tile's loads fit in registers.
struct TileKernels {
    int tx, ty, precision;
    const void *pk, *diag, *seg;
};

// TI: intensity storage type, TC: arithmetic of the corner solve
template <typename TI, typename TC, int TX, int TY>
TileKernels tk(int precision)
{
    return TileKernels{TX, TY, precision, reinterpret_cast<const void *>(&sweep_pk<TI, TC, TX, TY>),
                       reinterpret_cast<const void *>(&sweep_diag<TI, TC, TX, TY>),
                       reinterpret_cast<const void *>(&sweep_seg<TI, TC, TX, TY>)};
}

const TileKernels *find_tile_kernels(int precision, int tx, int ty)
{
    static const TileKernels table[] = {
        tk<float, float, 2, 2>(1), tk<float, float, 2, 4>(1), tk<float, float, 4, 2>(1),
        tk<float, float, 4, 4>(1), tk<float, float, 8, 2>(1), tk<float, float, 2, 8>(1),
        tk<double, double, 2, 2>(0), tk<double, double, 2, 4>(0), tk<double, double, 4, 2>(0),
        tk<float, double, 2, 2>(2), tk<float, double, 2, 4>(2), tk<float, double, 4, 2>(2),
    };
    for (const TileKernels &t : table)
        if (t.tx == tx && t.ty == ty && t.precision == precision) return &t;
    return nullptr;
}

}  // namespace

// Host-side state of the tile engine of one context.
struct PkState {
    PkProb *probs = nullptr;     // [2]: table for I_crn_old in buffer 0 / buffer 1
    int4 *items = nullptr;
    int nitems = 0;
    int *queue = nullptr;        // [0] counter, [1] epoch, [2] error word
    int *flags = nullptr;
    size_t lds = 0;
    int grid = 0;
    const TileKernels *k = nullptr;
    const void *kern = nullptr;  // the persistent kernel: tile items or column segments
    PkProb host[2];              // host copies of the two tables
};

void pk_destroy(Ctx &c)
{
    delete c.pk;
    c.pk = nullptr;
}

bool pk_supported(const Ctx &c) { return find_tile_kernels(c.precision, c.sg.TX, c.sg.TY) != nullptr; }

// workgroups that fit on the device at once (the persistent grid), at most nitems
static int resident_grid(const void *kern, size_t lds, int nitems)
{
    int dev = 0, ncu = 0, per_cu = 0;
    TRT_CHECK(hipGetDevice(&dev));
    TRT_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    TRT_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, PK_THREADS, lds));
    if (per_cu < 1) throw Error("sweep tile kernel does not fit on a compute unit");
    return std::max(1, std::min(nitems, per_cu * ncu));
}

// Work-list entries of diagonal d of one problem: tiles (txl, tyl = d - txl), or
// with segs column segments (txl, s = d - txl) of seg tiles.
static void diag_items(std::vector<int4> &items, int prob, int base, const SweepGeom &sg, int seg, int nseg, int d,
                       bool segs)
{
    const int nrow = segs ? nseg : sg.NTY;
    const int lo = std::max(0, d - (nrow - 1)), hi = std::min(d, sg.NTX - 1);
    for (int txl = lo; txl <= hi; ++txl) {
        const int r = d - txl;
        const int tyl = segs ? r * seg : r;
        for (int q = 0; q < 4; ++q)
            for (int gb = 0; gb < sg.nblk_g; ++gb) {
                const int flag = base + ((q * sg.nblk_g + gb) * sg.NTY + tyl) * sg.NTX + txl;
                items.push_back(make_int4(prob, flag, txl | (r << 16), q | (gb << 8)));
            }
    }
}

static int num_diags(const SweepGeom &sg, int nseg, bool segs) { return sg.NTX + (segs ? nseg : sg.NTY) - 1; }

// Tiles per column segment: the longest power of two that still leaves about
// eight work items per resident workgroup (long segments keep more H faces in
// registers; enough items keep the device busy).
static int seg_length(const SweepGeom &sg, int grid)
{
    if (const char *v = getenv("TRT_SEG_LEN")) {
        const int s = atoi(v);
        if (s >= 1) return std::min(s, sg.NTY);
    }
    int seg = 1;
    while (seg * 2 <= sg.NTY &&
           (long long)sg.NTX * ((sg.NTY + seg * 2 - 1) / (seg * 2)) * 4 * sg.nblk_g >= 8LL * grid)
        seg *= 2;
    return seg;
}

// Flag checks a dependency wait makes before it gives up and sets the error word
// (TRT_SPIN_LIMIT overrides it, e.g. 0 to exercise the error path: every wait then
// ends at its first check, every item still runs, and the sweep reports the error).
static int spin_limit()
{
    if (const char *v = getenv("TRT_SPIN_LIMIT")) return std::max(0, atoi(v));
    return 1 << 20;
}

// Doubles of LDS the cross-wave moment sums need ahead of the staging area: none when
// the tile's pairs fit MAX_OVERLAP_PASSES passes and the staging area can hold their
// sums once read; else one pass's worth.
constexpr int XS_WORDS = 3 * NMOM * 64;
static int xs_words(int npair, size_t staging_bytes)
{
    const int npass = (npair + 63) / 64;
    return (npass <= MAX_OVERLAP_PASSES && staging_bytes >= (size_t)npass * XS_WORDS * 8) ? 0 : XS_WORDS;
}

static size_t staging_bytes(int tx, int ty, int precision)
{
    return (size_t)3 * tx * ty * PK_THREADS * (precision ? 4 : 8);
}

// The moment weights of every direction, w, w ox^2, w oy^2, w ox oy, w ox, w oy, over the
// factor of the stored intensities, in device memory (read by scalar loads).
static const double *moment_table(Ctx &c)
{
    std::vector<double> w(c.nm), ox(c.nm), oy(c.nm), tab((size_t)c.nm * 6);
    // on the context's stream, after its uploads of the quadrature
    TRT_CHECK(hipMemcpyAsync(w.data(), c.w, c.nm * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    TRT_CHECK(hipMemcpyAsync(ox.data(), c.ox, c.nm * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    TRT_CHECK(hipMemcpyAsync(oy.data(), c.oy, c.nm * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    TRT_CHECK(hipStreamSynchronize(c.stream));
    const double isc = c.precision ? 1.0 / kIScale32 : 1.0;
    for (int m = 0; m < c.nm; ++m) {
        double *t = &tab[(size_t)m * 6];
        t[0] = w[m] * isc;
        t[1] = w[m] * ox[m] * ox[m] * isc;
        t[2] = w[m] * oy[m] * oy[m] * isc;
        t[3] = w[m] * ox[m] * oy[m] * isc;
        t[4] = w[m] * ox[m] * isc;
        t[5] = w[m] * oy[m] * isc;
    }
    double *d = static_cast<double *>(c.dalloc(tab.size() * sizeof(double)));
    TRT_CHECK(hipMemcpyAsync(d, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
    TRT_CHECK(hipStreamSynchronize(c.stream));
    return d;
}

static void pk_setup(Ctx &c)
{
    const SweepGeom &sg = c.sg;
    PkState *pk = new PkState();
    c.pk = pk;
    pk->k = find_tile_kernels(c.precision, sg.TX, sg.TY);
    if (!pk->k) throw Error("no tile kernel for this tile shape and precision");
    PkProb P{};
    P.nx = c.nx; P.ny = c.ny; P.ng = c.ng; P.K = sg.K; P.gblk = sg.gblk; P.NTX = sg.NTX; P.NTY = sg.NTY;
    P.npair = sg.TX * sg.TY * sg.gblk;
    P.kw = (sg.K + 3) / 4;
    const size_t stb = staging_bytes(sg.TX, sg.TY, c.precision);
    P.xsw = xs_words(P.npair, stb);
    P.ctab = moment_table(c);
    P.inv_cdt = 1.0 / (c.p.c * c.p.dt);
    P.eps = 1e-25 / c.p.comp_unit;
    P.dx = c.dx; P.dy = c.dy; P.ox = c.ox; P.oy = c.oy; P.w = c.w;
    P.sc = c.precision ? static_cast<const void *>(c.scf) : static_cast<const void *>(c.sc);
    P.bcin = c.bcin;
    P.XV = c.XV; P.XH = c.XH;
    P.Part = c.Part; P.PVb = c.PVb; P.PHb = c.PHb;
    P.PP = c.pl.PP; P.nblk_g = c.pl.nblk_g;
    P.NLr = rec_lanes(c.ng * sg.K);
    P.tx = sg.TX; P.ty = sg.TY;
    P.refl = c.refl; P.BV = c.BV; P.BH = c.BH;
    pk->lds = (size_t)P.xsw * 8 + stb;
    // The default form (sweep_mode 0) takes column segments of 4 or 2 tiles when their
    // wavefront (items of one diagonal) still offers twice the resident workgroups, as
    // batches do: on large grids segments save face traffic, on small ones the segment's
    // serial tiles would lengthen the critical path.  Results are bitwise the same.
    int auto_seg = 0;
    if (c.p.sweep_mode == 0 && c.sweep_mode == 2 && !getenv("TRT_SEG_LEN")) {
        const long long g2 = 2LL * resident_grid(pk->k->seg, pk->lds, 1 << 30);
        for (int s = 4; s >= 2 && auto_seg == 0; s /= 2)
            if ((long long)std::min(sg.NTX, (sg.NTY + s - 1) / s) * 4 * sg.nblk_g >= g2) auto_seg = s;
        if (auto_seg) c.sweep_mode = 4;
    }
    const bool segs = c.sweep_mode == 4;
    const int full_grid = resident_grid(segs ? pk->k->seg : pk->k->pk, pk->lds, 1 << 30);
    P.seg = segs ? (auto_seg ? std::min(auto_seg, sg.NTY) : seg_length(sg, full_grid)) : 1;
    P.nseg = (sg.NTY + P.seg - 1) / P.seg;
    PkProb tab[2] = {P, P};
    tab[0].Iold = c.Icrn[0]; tab[0].Inew = c.Icrn[1];
    tab[1].Iold = c.Icrn[1]; tab[1].Inew = c.Icrn[0];
    pk->host[0] = tab[0];
    pk->host[1] = tab[1];
    // every setup write goes on the context's stream, ahead of the sweeps that depend on it
    // (the stream does not wait for work on the null stream)
    pk->probs = static_cast<PkProb *>(c.dalloc(sizeof(tab)));
    TRT_CHECK(hipMemcpyAsync(pk->probs, tab, sizeof(tab), hipMemcpyHostToDevice, c.stream));

    // items in tile-diagonal (or segment-diagonal) order
    std::vector<int4> items;
    const int nchain = 4 * sg.nblk_g;
    items.reserve((size_t)nchain * sg.NTX * sg.NTY);
    for (int d = 0; d < num_diags(sg, P.nseg, segs); ++d) diag_items(items, 0, 0, sg, P.seg, P.nseg, d, segs);
    pk->nitems = (int)items.size();
    pk->items = static_cast<int4 *>(c.dalloc(items.size() * sizeof(int4)));
    TRT_CHECK(hipMemcpyAsync(pk->items, items.data(), items.size() * sizeof(int4), hipMemcpyHostToDevice, c.stream));
    pk->flags = static_cast<int *>(c.dalloc((size_t)nchain * sg.NTX * sg.NTY * sizeof(int)));
    TRT_CHECK(hipMemsetAsync(pk->flags, 0, (size_t)nchain * sg.NTX * sg.NTY * sizeof(int), c.stream));
    pk->queue = static_cast<int *>(c.dalloc(4 * sizeof(int)));
    TRT_CHECK(hipMemsetAsync(pk->queue, 0, 4 * sizeof(int), c.stream));
    TRT_CHECK(hipStreamSynchronize(c.stream));      // the host arrays above go out of scope
    pk->kern = segs ? pk->k->seg : pk->k->pk;
    pk->grid = std::min(full_grid, pk->nitems);
}

// boundary-face moments of nprob problems (table tab), the largest with nbt faces x groups x quadrants
static void boundary_moments(int precision, const PkProb *tab, int nprob, int nbt, hipStream_t s)
{
    const dim3 grid((nbt + 255) / 256, nprob);
    if (precision)
        hipLaunchKernelGGL(pk_boundary_moments<float>, grid, dim3(256), 0, s, tab);
    else
        hipLaunchKernelGGL(pk_boundary_moments<double>, grid, dim3(256), 0, s, tab);
}

static void boundary_moments(Ctx &c, const PkProb *tab)
{
    boundary_moments(c.precision, tab, 1, 4 * c.ng * (c.nx + c.ny), c.stream);
}

// values per boundary slice of a V side and of an H side
static void slice_sizes(const Ctx &c, size_t &nv, size_t &nh)
{
    const size_t NR = (size_t)rec_lanes(c.ng * c.sg.K);
    nv = (size_t)c.sg.NTY * NR * 2 * c.sg.TY;
    nh = (size_t)c.sg.NTX * NR * 2 * c.sg.TX;
}

void rt_bc_update(Ctx &c)
{
    if (!c.refl) return;
    size_t nv, nh;
    slice_sizes(c, nv, nh);
    const dim3 grid((unsigned)std::min<size_t>(1024, (std::max(nv, nh) + 255) / 256), 8);
    if (c.precision)
        hipLaunchKernelGGL(reflect_kernel<float>, grid, dim3(256), 0, c.stream, static_cast<float *>(c.BV),
                           static_cast<float *>(c.BH), nv, nh, c.refl);
    else
        hipLaunchKernelGGL(reflect_kernel<double>, grid, dim3(256), 0, c.stream, static_cast<double *>(c.BV),
                           static_cast<double *>(c.BH), nv, nh, c.refl);
}

void reflect_reset(Ctx &c)
{
    if (!c.refl) return;
    size_t nv, nh;
    slice_sizes(c, nv, nh);
    const dim3 grid((unsigned)std::min<size_t>(1024, (4 * std::max(nv, nh) + 255) / 256), 4);
    const int NL = c.ng * c.sg.K;
    if (c.precision)
        hipLaunchKernelGGL(reflect_reset_kernel<float>, grid, dim3(256), 0, c.stream, static_cast<float *>(c.BV),
                           static_cast<float *>(c.BH), c.bcin, nv, nh, NL, rec_lanes(NL), c.sg.K, c.ng, c.refl);
    else
        hipLaunchKernelGGL(reflect_reset_kernel<double>, grid, dim3(256), 0, c.stream, static_cast<double *>(c.BV),
                           static_cast<double *>(c.BH), c.bcin, nv, nh, NL, rec_lanes(NL), c.sg.K, c.ng, c.refl);
}

// Enqueues one sweep: work-list reset, boundary-face moments, persistent kernel.
// With ev, (*ev)[0] and (*ev)[1] bracket the persistent kernel alone.
void pk_sweep(Ctx &c, std::vector<hipEvent_t> *ev)
{
    if (!c.pk) pk_setup(c);
    PkState &pk = *c.pk;
    const PkProb *tab = pk.probs + c.iold;
    hipLaunchKernelGGL(pk_reset, dim3(1), dim3(64), 0, c.stream, pk.queue);
    boundary_moments(c, tab);
    PkArgs a{};
    a.probs = tab;
    a.items = pk.items;
    a.nitems = pk.nitems;
    a.queue = pk.queue;
    a.flags = pk.flags;
    a.spin_limit = spin_limit();
    void *args[] = {&a};
    if (ev) TRT_CHECK(hipEventRecord((*ev)[0], c.stream));
    TRT_CHECK(hipLaunchKernel(pk.kern, dim3(pk.grid), dim3(PK_THREADS), args, pk.lds, c.stream));
    if (ev) TRT_CHECK(hipEventRecord((*ev)[1], c.stream));
}

// The same sweep as one launch per tile diagonal.  With ev, (*ev)[2d], (*ev)[2d+1]
// bracket the launch of diagonal d.
void diag_sweep(Ctx &c, std::vector<hipEvent_t> *ev)
{
    if (!c.pk) pk_setup(c);
    PkState &pk = *c.pk;
    const SweepGeom &sg = c.sg;
    const PkProb *tab = pk.probs + c.iold;
    PkProb hp = pk.host[c.iold];
    boundary_moments(c, tab);
    for (int d = 0; d <= sg.NTX + sg.NTY - 2; ++d) {
        int tx_lo = std::max(0, d - (sg.NTY - 1));
        const int tx_hi = std::min(d, sg.NTX - 1);
        int nt_d = tx_hi - tx_lo + 1;
        int nblk_g = sg.nblk_g;
        void *args[] = {&hp, &d, &tx_lo, &nt_d, &nblk_g};
        if (ev) TRT_CHECK(hipEventRecord((*ev)[2 * d], c.stream));
        TRT_CHECK(hipLaunchKernel(pk.k->diag, dim3(4 * nt_d * nblk_g), dim3(PK_THREADS), args, pk.lds, c.stream));
        if (ev) TRT_CHECK(hipEventRecord((*ev)[2 * d + 1], c.stream));
    }
}

// Reads and clears the persistent sweep's error word (synchronises the stream).
void pk_check(Ctx &c)
{
    if (!c.pk) return;
    int e = 0;
    TRT_CHECK(hipMemcpyAsync(&e, c.pk->queue + 2, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    TRT_CHECK(hipStreamSynchronize(c.stream));
    if (e) {
        const int zero = 0;
        TRT_CHECK(hipMemcpyAsync(c.pk->queue + 2, &zero, sizeof(int), hipMemcpyHostToDevice, c.stream));
        TRT_CHECK(hipStreamSynchronize(c.stream));
        throw Error("persistent sweep: a dependency wait exceeded its spin limit");
    }
}

int pk_grid(const Ctx &c) { return c.pk ? c.pk->grid : 0; }

// ---------------------------------------------------------------------------
// Batched persistent sweep: one launch sweeps every problem of a batch.
// ---------------------------------------------------------------------------
struct PkBatch {
    std::vector<Ctx *> ctxs;
    const TileKernels *k = nullptr;
    int precision = 0, tx = 0, ty = 0;
    PkProb *probs = nullptr;        // device table, one descriptor per problem
    std::vector<PkProb> host;       // the table as last uploaded
    std::vector<int> up_iold;       // I_crn_old buffer of each problem at that upload
    PkProb *staging = nullptr;      // pinned copy of the table for the upload
    hipEvent_t uploaded = nullptr;  // the last upload has read the staging copy
    bool upload_pending = false;
    int4 *items = nullptr;
    int nitems = 0;
    int *queue = nullptr;           // [0] counter, [1] epoch, [2] error word
    int *flags = nullptr;
    size_t lds = 0;
    int grid = 0, nbt = 0;
    long long updates = 0;          // cell-angle-group updates per batched sweep
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr, t0 = nullptr, t1 = nullptr;
    std::vector<hipEvent_t> ready;  // per problem: its stream's work before the sweep
    std::vector<void *> owned;
    std::vector<size_t> owned_bytes;
    bool pooled = false;            // device blocks, staging and stream from the process caches
    bool segs = false;              // column-segment items
    int seg = 1;                    // column-segment length chosen for the batch
    std::vector<int> seg_of;        // tiles per segment of each problem
    const void *kern = nullptr;
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // whole Tgen_QDf steps as graphs, one per I_crn_old buffer signature of the members
    // (the handover swaps each member's two buffers every step, so a batch stepping
    // together alternates between two graphs)
    bool step_graphs_on = false;
    struct StepGraph {
        std::vector<int> sig;       // each member's I_crn_old buffer in the graph's step
        PkProb *table = nullptr;    // device descriptor table of that signature
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<StepGraph> step_graphs;
    hipEvent_t fork = nullptr, fork2 = nullptr;
    std::vector<hipEvent_t> joined;
};

static void *batch_alloc(PkBatch &b, size_t bytes)
{
    void *p = nullptr;
    size_t n = std::max<size_t>(bytes, 16);
    if (b.pooled) p = cache_alloc(n);
    else TRT_CHECK(hipMalloc(&p, n));
    b.owned.push_back(p);
    b.owned_bytes.push_back(n);
    return p;
}

void pk_batch_destroy(PkBatch *b)
{
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (auto &g : b->step_graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
    }
    for (hipEvent_t e : b->joined) (void)hipEventDestroy(e);
    for (hipEvent_t e : {b->fork, b->fork2})
        if (e) (void)hipEventDestroy(e);
    for (size_t k = 0; k < b->owned.size(); ++k) {
        if (b->pooled) cache_release(b->owned[k], b->owned_bytes[k]);
        else (void)hipFree(b->owned[k]);
    }
    for (hipEvent_t e : b->ready) (void)hipEventDestroy(e);
    for (hipEvent_t e : {b->done, b->t0, b->t1, b->uploaded})
        if (e) (void)hipEventDestroy(e);
    const size_t sw = b->ctxs.size() * sizeof(PkProb) / sizeof(double);
    if (b->staging) {
        if (b->pooled) cache_pinned_put(reinterpret_cast<double *>(b->staging), sw);
        else (void)hipHostFree(b->staging);
    }
    if (b->stream) {
        if (b->pooled) cache_stream_put(b->stream);
        else (void)hipStreamDestroy(b->stream);
    }
    delete b;
}

PkBatch *pk_batch_create(const std::vector<Ctx *> &ctxs, bool pooled, int seg_req)
{
    if (ctxs.empty()) throw Error("a batch needs at least one context");
    Ctx &c0 = *ctxs[0];
    for (Ctx *c : ctxs) {
        if (c->precision != c0.precision) throw Error("batched contexts must share the sweep precision");
        if (c->sweep_mode < 2) throw Error("batched contexts must use the tile engine (sweep_mode 0, 2, 3 or 4)");
        if (c->sg.TX != c0.sg.TX || c->sg.TY != c0.sg.TY) throw Error("batched contexts must share the tile shape");
    }
    for (size_t i = 0; i < ctxs.size(); ++i)
        for (size_t j = 0; j < i; ++j)
            if (ctxs[i] == ctxs[j]) throw Error("a context appears twice in the batch");
    PkBatch *b = new PkBatch();
    try {
        b->ctxs = ctxs;
        b->pooled = pooled;
        b->precision = c0.precision;
        b->tx = c0.sg.TX;
        b->ty = c0.sg.TY;
        b->k = find_tile_kernels(b->precision, b->tx, b->ty);
        if (!b->k) throw Error("no tile kernel for this tile shape and precision");
        const int n = (int)ctxs.size();
        b->host.resize(n);
        b->up_iold.assign(n, -1);
        // The batch picks its own item form (the results are bitwise the same): the
        // requested segment length, or by default column segments of up to 4 tiles when
        // the segment wavefronts of all members still offer twice the resident
        // workgroups, else single tiles.
        {
            const size_t stb0 = staging_bytes(b->tx, b->ty, b->precision);
            int xsw0 = 0;
            for (Ctx *c : ctxs) xsw0 = std::max(xsw0, xs_words(b->tx * b->ty * c->sg.gblk, stb0));
            const size_t lds0 = (size_t)xsw0 * 8 + stb0;
            const long long g2 = 2LL * resident_grid(b->k->seg, lds0, 1 << 30);
            int seg = 1;
            if (seg_req > 0) seg = seg_req;
            else if (const char *v = getenv("TRT_SEG_LEN")) seg = std::max(1, atoi(v));
            else
                for (int s = 4; s >= 2 && seg == 1; s /= 2) {
                    long long par = 0;
                    for (Ctx *c : ctxs)
                        par += (long long)std::min(c->sg.NTX, (c->sg.NTY + s - 1) / s) * 4 * c->sg.nblk_g;
                    if (par >= g2) seg = s;
                }
            b->segs = seg > 1;
            b->seg = seg;
            b->seg_of.resize(n);
            for (int i = 0; i < n; ++i) b->seg_of[i] = std::min(seg, ctxs[i]->sg.NTY);
        }
        std::vector<int> base(n + 1, 0);
        int maxd = 0, maxxs = 0;
        size_t expect = 0;
        for (int i = 0; i < n; ++i) {
            Ctx &c = *ctxs[i];
            if (!c.pk) pk_setup(c);
            const SweepGeom &sg = c.sg;
            base[i + 1] = base[i] + 4 * sg.nblk_g * sg.NTX * sg.NTY;
            const int nseg = (sg.NTY + b->seg_of[i] - 1) / b->seg_of[i];
            maxd = std::max(maxd, num_diags(sg, nseg, b->segs) - 1);
            expect += (size_t)4 * sg.nblk_g * sg.NTX * (b->segs ? nseg : sg.NTY);
            maxxs = std::max(maxxs, c.pk->host[0].xsw);
            b->nbt = std::max(b->nbt, 4 * c.ng * (c.nx + c.ny));
            b->updates += (long long)c.nx * c.ny * c.nm * c.ng;
        }
        // items of all problems, merged by tile (or segment) diagonal
        std::vector<int4> items;
        items.reserve((size_t)base[n]);
        for (int d = 0; d <= maxd; ++d)
            for (int i = 0; i < n; ++i) {
                const SweepGeom &sg = ctxs[i]->sg;
                const int seg = b->seg_of[i], nseg = (sg.NTY + seg - 1) / seg;
                if (d < num_diags(sg, nseg, b->segs)) diag_items(items, i, base[i], sg, seg, nseg, d, b->segs);
            }
        b->nitems = (int)items.size();
        if ((size_t)b->nitems != expect) throw Error("batched work list does not cover every tile");
        b->items = static_cast<int4 *>(batch_alloc(*b, items.size() * sizeof(int4)));
        b->flags = static_cast<int *>(batch_alloc(*b, (size_t)base[n] * sizeof(int)));
        b->queue = static_cast<int *>(batch_alloc(*b, 4 * sizeof(int)));
        b->probs = static_cast<PkProb *>(batch_alloc(*b, (size_t)n * sizeof(PkProb)));
        b->lds = (size_t)maxxs * 8 + staging_bytes(b->tx, b->ty, b->precision);
        b->kern = b->segs ? b->k->seg : b->k->pk;
        b->grid = resident_grid(b->kern, b->lds, b->nitems);
        if (pooled) b->stream = cache_stream_get();
        else TRT_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        // the work list, flags and counter on the batch stream, ahead of its sweeps
        TRT_CHECK(hipMemcpyAsync(b->items, items.data(), items.size() * sizeof(int4), hipMemcpyHostToDevice,
                                 b->stream));
        TRT_CHECK(hipMemsetAsync(b->flags, 0, (size_t)base[n] * sizeof(int), b->stream));
        TRT_CHECK(hipMemsetAsync(b->queue, 0, 4 * sizeof(int), b->stream));
        TRT_CHECK(hipStreamSynchronize(b->stream));
        TRT_CHECK(hipEventCreateWithFlags(&b->done, hipEventDisableTiming));
        TRT_CHECK(hipEventCreateWithFlags(&b->uploaded, hipEventDisableTiming));
        static_assert(sizeof(PkProb) % sizeof(double) == 0, "table staged as doubles");
        const size_t sw = (size_t)n * sizeof(PkProb) / sizeof(double);
        if (pooled) b->staging = reinterpret_cast<PkProb *>(cache_pinned_get(sw));
        else TRT_CHECK(hipHostMalloc(reinterpret_cast<void **>(&b->staging), sw * sizeof(double), hipHostMallocDefault));
        TRT_CHECK(hipEventCreate(&b->t0));
        TRT_CHECK(hipEventCreate(&b->t1));
        b->ready.resize(n, nullptr);
        for (auto &e : b->ready) TRT_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    } catch (...) {
        pk_batch_destroy(b);
        throw;
    }
    return b;
}

static void batch_enqueue(PkBatch &b, const PkProb *table = nullptr)
{
    if (!table) table = b.probs;
    hipLaunchKernelGGL(pk_reset, dim3(1), dim3(64), 0, b.stream, b.queue);
    boundary_moments(b.precision, table, (int)b.ctxs.size(), b.nbt, b.stream);
    PkArgs a{};
    a.probs = table;
    a.items = b.items;
    a.nitems = b.nitems;
    a.queue = b.queue;
    a.flags = b.flags;
    a.spin_limit = spin_limit();
    void *args[] = {&a};
    TRT_CHECK(hipLaunchKernel(b.kern, dim3(b.grid), dim3(PK_THREADS), args, b.lds, b.stream));
}

// One batched sweep: every problem's TRANSPORT_SCB (I_crn from I_crn_old) with
// its moments, after the work already enqueued on each problem's stream; each
// problem's stream waits for the sweep.  Events t0 (mark & 1) and t1 (mark & 2)
// are recorded on the batch stream before and after the sweep.
void pk_batch_sweep(PkBatch &b, int mark, const std::vector<char> *sweep)
{
    const int n = (int)b.ctxs.size();
    if (sweep && (int)sweep->size() != n) throw Error("batched sweep: the member mask has the wrong length");
    auto swept = [&](int i) { return !sweep || (*sweep)[i]; };
    bool dirty = false;
    for (int i = 0; i < n; ++i) {
        Ctx &c = *b.ctxs[i];
        const int skip = swept(i) ? 0 : 1;
        if (b.up_iold[i] != c.iold || b.host[i].skip != skip) {
            b.host[i] = c.pk->host[c.iold];
            b.host[i].seg = b.seg_of[i];
            b.host[i].nseg = (b.host[i].NTY + b.seg_of[i] - 1) / b.seg_of[i];
            b.host[i].skip = skip;
            b.up_iold[i] = c.iold;
            dirty = true;
        }
    }
    if (dirty) {
        if (b.upload_pending) TRT_CHECK(hipEventSynchronize(b.uploaded));   // staging copy free again
        std::copy(b.host.begin(), b.host.end(), b.staging);
        TRT_CHECK(hipMemcpyAsync(b.probs, b.staging, (size_t)n * sizeof(PkProb), hipMemcpyHostToDevice, b.stream));
        TRT_CHECK(hipEventRecord(b.uploaded, b.stream));
        b.upload_pending = true;
    }
    for (int i = 0; i < n; ++i) {
        if (!swept(i)) continue;
        TRT_CHECK(hipEventRecord(b.ready[i], b.ctxs[i]->stream));
        TRT_CHECK(hipStreamWaitEvent(b.stream, b.ready[i], 0));
    }
    if (mark & 1) TRT_CHECK(hipEventRecord(b.t0, b.stream));
    if (b.use_graph) {
        if (!b.exec) {
            TRT_CHECK(hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal));
            batch_enqueue(b);
            TRT_CHECK(hipStreamEndCapture(b.stream, &b.graph));
            TRT_CHECK(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
        }
        TRT_CHECK(hipGraphLaunch(b.exec, b.stream));
    } else {
        batch_enqueue(b);
    }
    if (mark & 2) TRT_CHECK(hipEventRecord(b.t1, b.stream));
    TRT_CHECK(hipGetLastError());
    TRT_CHECK(hipEventRecord(b.done, b.stream));
    for (int i = 0; i < n; ++i) {
        if (!swept(i)) continue;
        Ctx &c = *b.ctxs[i];
        TRT_CHECK(hipStreamWaitEvent(c.stream, b.done, 0));
        c.icur = 1 - c.iold;
        c.sweeps += 1;
    }
}

void pk_batch_use_graph(PkBatch &b, int mode)
{
    b.use_graph = mode == 1;
    b.step_graphs_on = mode == 2;
}

bool pk_batch_step_graphs(const PkBatch &b) { return b.step_graphs_on; }

// One Tgen_QDf step of every member (MATERIAL_UPDATE, I_crn_old = I_crn, the batched
// TRANSPORT_SCB, COLLAPSE_INTENSITIES + fg_Calc, Cg_Calc) as one graph launch on the batch
// stream.  The graph forks to each member's stream for its MATERIAL_UPDATE, joins for the
// persistent sweep, and forks again for the moments, QD factors and C_g; it is captured
// on first use for each signature of the members' I_crn_old buffers and replayed after.
// The captured work is kernels only (no memset or copy nodes).  Events t0 and t1 bracket
// the whole step.
void pk_batch_tgen_step_graph(PkBatch &b)
{
    const int n = (int)b.ctxs.size();
    std::vector<int> sig(n);
    for (int i = 0; i < n; ++i) sig[i] = b.ctxs[i]->icur;   // the buffer the handover makes I_crn_old
    PkBatch::StepGraph *sg = nullptr;
    for (auto &g : b.step_graphs)
        if (g.sig == sig) sg = &g;
    // the step starts after the work already enqueued on every member's stream
    for (int i = 0; i < n; ++i) {
        TRT_CHECK(hipEventRecord(b.ready[i], b.ctxs[i]->stream));
        TRT_CHECK(hipStreamWaitEvent(b.stream, b.ready[i], 0));
    }
    if (!sg) {
        if (!b.fork) {
            TRT_CHECK(hipEventCreateWithFlags(&b.fork, hipEventDisableTiming));
            TRT_CHECK(hipEventCreateWithFlags(&b.fork2, hipEventDisableTiming));
            b.joined.resize(n, nullptr);
            for (auto &e : b.joined) TRT_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
        PkBatch::StepGraph g;
        g.sig = sig;
        std::vector<PkProb> host(n);
        for (int i = 0; i < n; ++i) {
            host[i] = b.ctxs[i]->pk->host[sig[i]];
            host[i].seg = b.seg_of[i];
            host[i].nseg = (host[i].NTY + b.seg_of[i] - 1) / b.seg_of[i];
        }
        g.table = static_cast<PkProb *>(batch_alloc(b, (size_t)n * sizeof(PkProb)));
        TRT_CHECK(hipMemcpyAsync(g.table, host.data(), (size_t)n * sizeof(PkProb), hipMemcpyHostToDevice, b.stream));
        TRT_CHECK(hipStreamSynchronize(b.stream));
        TRT_CHECK(hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal));
        TRT_CHECK(hipEventRecord(b.fork, b.stream));
        for (int i = 0; i < n; ++i) {
            Ctx &c = *b.ctxs[i];
            TRT_CHECK(hipStreamWaitEvent(c.stream, b.fork, 0));
            material_update(c);
            TRT_CHECK(hipEventRecord(b.joined[i], c.stream));
            TRT_CHECK(hipStreamWaitEvent(b.stream, b.joined[i], 0));
        }
        batch_enqueue(b, g.table);
        TRT_CHECK(hipEventRecord(b.fork2, b.stream));
        for (int i = 0; i < n; ++i) {
            Ctx &c = *b.ctxs[i];
            TRT_CHECK(hipStreamWaitEvent(c.stream, b.fork2, 0));
            tgen_post(c);
            TRT_CHECK(hipEventRecord(b.joined[i], c.stream));
            TRT_CHECK(hipStreamWaitEvent(b.stream, b.joined[i], 0));
        }
        TRT_CHECK(hipStreamEndCapture(b.stream, &g.graph));
        TRT_CHECK(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
        b.step_graphs.push_back(g);
        sg = &b.step_graphs.back();
    }
    for (Ctx *c : b.ctxs) {        // the host side of the step (tgen_pre without its kernel)
        c->time += c->p.dt;
        c->iold = c->icur;
    }
    TRT_CHECK(hipEventRecord(b.t0, b.stream));
    TRT_CHECK(hipGraphLaunch(sg->exec, b.stream));
    TRT_CHECK(hipEventRecord(b.t1, b.stream));
    TRT_CHECK(hipGetLastError());
    TRT_CHECK(hipEventRecord(b.done, b.stream));
    for (int i = 0; i < n; ++i) {
        Ctx &c = *b.ctxs[i];
        TRT_CHECK(hipStreamWaitEvent(c.stream, b.done, 0));
        c.icur = 1 - c.iold;
        c.sweeps += 1;
    }
}

// waits for the batch stream; throws if a dependency wait of the last sweeps timed out
void pk_batch_check(PkBatch &b)
{
    int e = 0;
    TRT_CHECK(hipMemcpyAsync(&e, b.queue + 2, sizeof(int), hipMemcpyDeviceToHost, b.stream));
    TRT_CHECK(hipStreamSynchronize(b.stream));
    if (e) {
        const int zero = 0;
        TRT_CHECK(hipMemcpyAsync(b.queue + 2, &zero, sizeof(int), hipMemcpyHostToDevice, b.stream));
        TRT_CHECK(hipStreamSynchronize(b.stream));
        throw Error("batched sweep: a dependency wait exceeded its spin limit");
    }
}

double pk_batch_elapsed_ms(PkBatch &b)
{
    float ms = 0.f;
    TRT_CHECK(hipEventSynchronize(b.t1));
    TRT_CHECK(hipEventElapsedTime(&ms, b.t0, b.t1));
    return ms;
}

int pk_batch_seg(const PkBatch &b) { return b.seg; }

void pk_batch_info(const PkBatch &b, long long *updates, int *grid, int *nitems)
{
    if (updates) *updates = b.updates;
    if (grid) *grid = b.grid;
    if (nitems) *nitems = b.nitems;
}

const std::vector<Ctx *> &pk_batch_ctxs(const PkBatch &b) { return b.ctxs; }

}  // namespace trt
