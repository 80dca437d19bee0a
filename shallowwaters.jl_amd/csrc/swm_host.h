// swm_host.h -- host-side state of the gfx950 ShallowWaters time step: contexts, their
// device arena and slots, ensembles and the adjoint's state, shared by the C ABI
// (swm.hip) and the per-format launch units (swm_f64.hip, swm_f32.hip, swm_f16.hip,
// swm_bf16.hip).
//
// The driver replays the loop of time_integration.jl:45-243 as kernel launches on one HIP
// stream.  The reference's arrays u, u0, u1 (and v, η alike) become four device slots per
// field.  Slots take roles instead of being copied (copyto! at :49-51, :151-154, :191-193,
// :240-242 become a change of role): each stage writes into a slot that holds neither Prog
// nor a value still to be read.  Dissipation writes the new Prog into a free slot, so the
// previous Prog stays intact until the step is accepted: when NaNs are detected the loop
// stops before the copy-back exactly as the reference does (time_integration.jl:235-237).
//
// Number formats.  T = Tprog in {Float64, Float32}, or T in {Float16, BFloat16} with
// Tprog = Float32 (the classic RK scheme).  In the mixed case the stage input u1 between
// RK stages is held in T: the reference converts it to T before rhs! and uses it for
// nothing else (time_integration.jl:57-64), so two extra slots of T storage replace the
// Tprog u1 without changing a bit of the result.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <string>
#include <vector>

#include "../../include/swm.h"
#include "adj_core.h"
#include "swm_common.h"

using namespace swm;

#define HIPCHK(expr)                                                                                \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return ctx->fail(SWM_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

namespace swmh {

struct Field {             // one padded device array, carved from the context's arena
    void *base = nullptr;  // start of the field
    char *p0 = nullptr;    // element (0,0)
    uint32_t origin = 0;   // byte offset of base in the arena
    int es = 4;            // element size of its storage format
    uint32_t ref() const { return origin; }
};

struct Slots {             // where a set (u, v, η) lives: slot indices, in the T pool if t
    int uv = -1, e = -1;
    bool t = false;
};

// host rounding of a Float64 value to a number format (Julia's T(x), one rounding)
inline float f32_round_odd_host(double x) {
    float f = (float)x;
    if ((double)f != x && std::isfinite(f)) {
        uint32_t b;
        std::memcpy(&b, &f, 4);
        if (std::fabs((double)f) > std::fabs(x)) b -= 1u;
        b |= 1u;
        std::memcpy(&f, &b, 4);
    }
    return f;
}
inline double round_host(double x, int fmt) {
    if (fmt == SWM_F64 || std::isnan(x)) return x;
    if (fmt == SWM_F32) return (double)(float)x;
    const float f = f32_round_odd_host(x);
    if (fmt == SWM_BF16) {
        uint32_t b;
        std::memcpy(&b, &f, 4);
        b += 0x7FFFu + ((b >> 16) & 1u);
        b &= 0xFFFF0000u;
        float r;
        std::memcpy(&r, &b, 4);
        return r;
    }
    // binary16: quantum 2^-24 below 2^-14, 2^(e-10) above; overflow to infinity
    const double a = std::fabs((double)f);
    if (std::isinf(f) || a >= 65520.0) return std::copysign(INFINITY, x);
    int e;
    std::frexp(a, &e);                             // a = m * 2^e, m in [0.5, 1)
    const int q = std::max(e - 11, -24);           // log2 of the quantum
    const double r = std::nearbyint(std::ldexp(a, -q));
    return std::copysign(std::ldexp(r, q), x);
}

}  // namespace swmh

using namespace swmh;

// Kernels of a step recorded instead of launched (ensemble planning): kind 0 single stage
// (mode ma), 1 fused pair (ma, mb), 2 dissipation, 3 semi-implicit continuity, 4 / 5 periodic
// pair strips clear of the seam / at it, 6 update of the kept q and KEu+KEv (ma: bit 0 q
// too, bit 1 from the initial state); args bytes.
struct RecItem {
    int kind, ma, mb;
    std::vector<char> args;
};
struct Recorder {
    std::vector<RecItem> items;
    void add(int kind, int ma, int mb, const void *a, size_t n) {
        RecItem it{kind, ma, mb, std::vector<char>((const char *)a, (const char *)a + n)};
        items.push_back(std::move(it));
    }
};

// In-step kernel timing (swm_time_steps): HIP events recorded around each launch of the
// step's kernels, per kind of kernel, begin and end alternately.
enum { KT_FUSED12 = 0, KT_FUSED34, KT_DISSIPATION, KT_TRACER, KT_STEP, KT_STAGE_FIRST, KT_STAGE_MID, KT_STAGE_LAST,
       KT_KINDS };
struct StepTrace {
    std::vector<hipEvent_t> ev[KT_KINDS];
};

struct swm_ctx {
    swm_config c;
    Recorder *rec = nullptr;   // set while an ensemble plans a step
    StepTrace *strace = nullptr;   // set while swm_time_steps runs
    std::string err;
    int dev = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;   // periodic: the seam strips of a fused pair, beside the others
    hipEvent_t fork = nullptr, join = nullptr;
    int fmt = SWM_F32;     // format bundle: SWM_F64, SWM_F32 (T = Tprog), SWM_F16, SWM_BF16 (T, Tprog = F32),
                           // FMT_F16A, FMT_BF16A (T's values, float arithmetic, Tprog = F32)
    int es = 4;            // element size of Tprog fields
    int est = 4;           // element size of T fields
    bool mixed = false;
    int P = 0, rows = 0;   // pitch (elements) and allocated rows of 2-D fields
    size_t elems = 0;
    int ep = 0, nux = 0, nuy = 0, nvx = 0, nvy = 0;
    bool fuse = true;      // run consecutive RK stages as fused pairs where supported
    int L = 64;            // rows per wave segment of the stage kernels
    int Ld = 64;           // rows per wave segment of the dissipation kernel
    int Lf = 64;           // rows per wave segment of the fused stage pairs
    Field u[12], v[12], e[12], sst[2];                       // Tprog (slots 4..11: the adjoint's taped step)
    Field tu[2], tv[2], te[2];                               // T: u1 between RK stages (mixed formats)
    Field um, vm, H, Fx, Fy, fq, eta_ref, Feta, qend;         // T
    Field f_u, f_v;                                          // T: f on the u-, v-grid (linear dynamics)
    Field qs, ks;                                            // T: kept q and KEu+KEv (nstep_advcor > 0)
    bool advcor_pending = false;   // a new state: q, KEu+KEv to be formed from it before its first stage
    Field cu, cv, ce;                                        // Kahan compensations (compensated RK), Tprog
    double sc = 1.0, sci = 1.0;                              // u, v carried as sc·u, sc·v
    int Fx_sx = 1, Fx_sy = 0, Fy_sx = 1, Fy_sy = 0, fq_sx = 1, fq_sy = 0, fu_sx = 1, fu_sy = 0, fv_sx = 1, fv_sy = 0;
    bool have_fuv = false;   // swm_set_static_linear done
    bool seam_ok = false;    // periodic: the fused stage pairs may wrap around the seam (set_static)
    int pu = 0, pe = 0;    // slots holding Prog u,v and Prog η
    int rsst = 0;
    Slots rk0, rk1;        // u0 / u1 after swm_stages
    int *dflags = nullptr; // [0] NaN flag, [1] tracer index error
    double *dpart = nullptr;
    bool have_static = false, have_prog = false;
    int nslots = 4;                // Tprog slot sets in use (6 to 11 once the adjoint is set up)
    int ens_refs = 0;              // ensembles holding this context (their plans point into the arena)
    struct AdjState *adj = nullptr;
    bool adj_fx = false;           // pullbacks accumulate the cotangent of Fx (swm_adj_track_Fx)
    std::vector<void *> allocs;
    void *arena = nullptr;         // every field of the context; kernels address it through one buffer descriptor
    size_t arena_bytes = 0, arena_used = 0;
    ArenaRef aref() const { return ArenaRef{arena, (uint32_t)arena_bytes}; }

    int fail(int code, const char *fmt_, ...) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt_);
        vsnprintf(buf, sizeof buf, fmt_, ap);
        va_end(ap);
        err = buf;
        return code;
    }
    template <class S> S *at(const Field &f) const { return (S *)f.p0; }
    const Field &fu(Slots s) const { return s.t ? tu[s.uv] : u[s.uv]; }
    const Field &fv(Slots s) const { return s.t ? tv[s.uv] : v[s.uv]; }
    const Field &fe(Slots s) const { return s.t ? te[s.e] : e[s.e]; }
};

namespace swmh {

// the lowest slot not in `busy` (four slots always leave one free: at most three are live)
inline int pick(std::initializer_list<int> busy) {
    for (int s = 0; s < 4; ++s) {
        bool used = false;
        for (int b : busy) used |= b == s;
        if (!used) return s;
    }
    return -1;
}

// Carve a field from the arena (zeroed at creation).  A 2-D field has P*rows elements of
// size es; a 1-D per-row field `elems_1d`.
inline int alloc_field(swm_ctx *ctx, Field &f, size_t elems_1d, bool two_d, int es) {
    const size_t n = two_d ? ctx->elems : elems_1d;
    const size_t at = (ctx->arena_used + 255) & ~(size_t)255;
    if (at + n * es > ctx->arena_bytes) return ctx->fail(SWM_E_ARG, "arena exhausted");
    ctx->arena_used = at + n * es;
    f.base = (char *)ctx->arena + at;
    f.origin = (uint32_t)at;
    f.es = es;
    f.p0 = (char *)f.base + (two_d ? ((size_t)OX + (size_t)OY * ctx->P) : (size_t)OY) * es;
    return SWM_OK;
}

// the 1-D per-row layout of a field carved as 2-D (x-invariant static fields)
inline void as_rows(Field &f) { f.p0 = (char *)f.base + (size_t)OY * f.es; }

// element (x, y) of a 2-D field, as a byte pointer
inline char *elem(const swm_ctx *ctx, const Field &f, int x, int y) {
    return f.p0 + ((ptrdiff_t)x + (ptrdiff_t)y * ctx->P) * f.es;
}

// copy a Julia (m, n) host array whose element [1,1] sits at (x0, y0)
inline int put2d(swm_ctx *ctx, const Field &f, const void *host, int m, int n, int x0, int y0) {
    HIPCHK(hipMemcpy2DAsync(elem(ctx, f, x0, y0), (size_t)ctx->P * f.es, host, (size_t)m * f.es, (size_t)m * f.es, n,
                            hipMemcpyHostToDevice, ctx->stream));
    return SWM_OK;
}
inline int get2d(swm_ctx *ctx, const Field &f, void *host, int m, int n, int x0, int y0) {
    HIPCHK(hipMemcpy2DAsync(host, (size_t)m * f.es, elem(ctx, f, x0, y0), (size_t)ctx->P * f.es, (size_t)m * f.es, n,
                            hipMemcpyDeviceToHost, ctx->stream));
    return SWM_OK;
}

// a T-format field downloaded into a Tprog (Float32) host array: exact widening
inline int get2d_widen(swm_ctx *ctx, const Field &f, void *host, int m, int n, int x0, int y0) {
    std::vector<uint16_t> tmp((size_t)m * n);
    int rc = get2d(ctx, f, tmp.data(), m, n, x0, y0);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    float *out = (float *)host;
    for (size_t k = 0; k < tmp.size(); ++k) {
        uint32_t b;
        if (ctx->c.T == SWM_BF16) b = (uint32_t)tmp[k] << 16;
        else {                                     // binary16 -> binary32
            const uint32_t h = tmp[k], s = (h >> 15) << 31, ex = (h >> 10) & 0x1f, mt = h & 0x3ff;
            if (ex == 0) {
                const float r = std::ldexp((float)mt, -24);
                std::memcpy(&b, &r, 4);
                b |= s;
            } else if (ex == 31) b = s | 0x7f800000u | (mt << 13);
            else b = s | ((ex + 112) << 23) | (mt << 13);
        }
        std::memcpy(&out[k], &b, 4);
    }
    return SWM_OK;
}

// A static field that does not vary along x is kept as one value per row (stride 0 in x).
inline int put_static(swm_ctx *ctx, Field &f, int *sx, int *sy, const void *host, int m, int n, int x0, int y0) {
    const char *h = (const char *)host;
    const int es = f.es;
    bool xinv = true;
    for (int j = 0; j < n && xinv; ++j)
        for (int i = 1; i < m && xinv; ++i)
            xinv = std::memcmp(h + ((size_t)j * m + i) * es, h + (size_t)j * m * es, es) == 0;
    if (xinv) {
        as_rows(f);
        std::vector<char> col((size_t)n * es);
        for (int j = 0; j < n; ++j) std::memcpy(&col[(size_t)j * es], h + (size_t)j * m * es, es);
        HIPCHK(hipMemcpyAsync(f.p0 + (ptrdiff_t)y0 * es, col.data(), col.size(), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        *sx = 0;
        *sy = 1;
        return SWM_OK;
    }
    f.p0 = (char *)f.base + ((size_t)OX + (size_t)OY * ctx->P) * es;
    *sx = 1;
    *sy = ctx->P;
    return put2d(ctx, f, host, m, n, x0, y0);
}

// blocks for nwaves wavefronts, padded to a multiple of the XCD count (wave_tile's remap)
inline int nblocks_waves(int nwaves) {
    const int nb = (nwaves + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK;
    return (nb + NXCD - 1) / NXCD * NXCD;
}

// the right-hand side's form: the PV scheme or the linear dynamics, and whether q and the
// kinetic part of the Bernoulli potential are kept between updates (nstep_advcor > 0 with the
// nonlinear dynamics, rhs.jl:36-38, time_integration.jl:208)
inline int adv_form(const swm_config &c) {
    if (c.dynamics == SWM_LINEAR) return ADV_LIN;
    const int s = c.adv_scheme == SWM_ADV_SADOURNY ? ADV_SAD : ADV_AH;
    return c.nstep_advcor > 0 ? (s | ADV_KEEP) : s;
}

// a T-format scalar the host already rounded (config values marked (T)) as the kernel's R
template <class R> R mkR(double d) {
    if constexpr (std::is_same_v<R, double> || std::is_same_v<R, float>) return (R)d;
    else {
        R r;
        r.v = (float)d;
        return r;
    }
}
}  // namespace swmh

// ====================================================================== adjoint
//
// Reverse mode of time_integration (classic RK, T = Tprog = Tcomm in {Float64, Float32}).
// swm_adj_forward runs the steps as swm_step does and keeps device snapshots of full states
// (u, v, η, sst and the tracer's um, vm) at the positions a binomial checkpointing schedule
// chooses for the given number of snapshots (Griewank's revolve: with s snapshots, n steps
// are reversed with the least repetition number r, n <= C(s+r, s)).  swm_adj_pullback
// reverses the steps last first: it restores the nearest snapshot, advances to the step,
// runs that step once more keeping its intermediate states (the taped step: every RK stage
// input, u0 before and after the dissipation, the tracer state it read), and applies the
// reverse step of adj_core.h.  The result is the gradient of <ȳ, y> with respect to the
// state given to swm_adj_forward (interior cells; the halo of the gradient is zero).

struct AdjState {
    swm::adj::AdjFields F{};       // adjoint fields (origins); roles rotate during a pullback
    swm::adj::AdjFields F0{};
    int La = 64;                   // rows per wave segment of the adjoint marches
    std::vector<char *> snap;      // snapshots: [0] the initial state, [1..s] the schedule's, [s+1] the final state
    char *snap_block = nullptr;    // one device allocation holding them all (kept while the layout repeats)
    std::vector<int64_t> at;       // step held by each schedule snapshot (-1: free)
    std::vector<char> snap_pending; // each snapshot's kept q and KEu+KEv still to be formed (nstep_advcor > 0)
    bool keep_init = false;        // the forward run formed them from its initial state
    size_t snap_bytes = 0;
    int *dmax = nullptr;           // the tracer adjoint's gather window (device, 2 ints)
    int64_t i0 = 0, t0 = 0, n = 0;
    int nsnap = 0;
    bool recorded = false;
    int64_t fwd_steps = 0;         // forward steps run by the last pullback (taped steps included)
    bool fx_valid = false;         // the last pullback accumulated the cotangent of Fx
    double sweep_ms = 0.0;         // device time of the last reverse sweep (HIP events)
    void release() {
        if (snap_block) (void)hipFree(snap_block);
        snap_block = nullptr;
        snap.clear();
        at.clear();
        if (dmax) (void)hipFree(dmax);
        dmax = nullptr;
    }
};

namespace swmh {

// the fields of a context, for re-basing after the arena moves
inline std::vector<Field *> all_fields(swm_ctx *ctx) {
    std::vector<Field *> f;
    for (int s = 0; s < 12; ++s) { f.push_back(&ctx->u[s]); f.push_back(&ctx->v[s]); f.push_back(&ctx->e[s]); }
    for (int s = 0; s < 2; ++s) {
        f.push_back(&ctx->sst[s]); f.push_back(&ctx->tu[s]); f.push_back(&ctx->tv[s]); f.push_back(&ctx->te[s]);
    }
    for (Field *g : {&ctx->um, &ctx->vm, &ctx->H, &ctx->Fx, &ctx->Fy, &ctx->fq, &ctx->eta_ref, &ctx->Feta, &ctx->qend,
                     &ctx->cu, &ctx->cv, &ctx->ce, &ctx->f_u, &ctx->f_v, &ctx->qs, &ctx->ks})
        f.push_back(g);
    return f;
}

// Enlarge the arena by `extra` bytes, keeping every field at its byte offset.
inline int grow_arena(swm_ctx *ctx, size_t extra) {
    const size_t nb = ctx->arena_bytes + extra;
    if (nb > 0xFFFF0000ull)
        return ctx->fail(SWM_E_UNSUPPORTED, "grid %dx%d with the adjoint needs %zu bytes of device fields; one context "
                         "addresses at most 4 GiB", ctx->c.nx, ctx->c.ny, nb);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    void *p = nullptr;
    if (hipMalloc(&p, nb) != hipSuccess) return ctx->fail(SWM_E_HIP, "hipMalloc of %zu bytes failed", nb);
    HIPCHK(hipMemsetAsync(p, 0, nb, ctx->stream));
    HIPCHK(hipMemcpyAsync(p, ctx->arena, ctx->arena_used, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    char *old = (char *)ctx->arena;
    for (Field *f : all_fields(ctx))
        if (f->base) {
            f->p0 = (char *)p + (f->p0 - old);
            f->base = (char *)p + f->origin;
        }
    for (auto &a : ctx->allocs)
        if (a == ctx->arena) a = p;
    (void)hipFree(ctx->arena);
    ctx->arena = p;
    ctx->arena_bytes = nb;
    return SWM_OK;
}

inline int64_t binom_beta(int s, int t) {               // C(s+t, s), saturating
    double b = 1.0;
    for (int k = 1; k <= s; ++k) b = b * (double)(t + k) / (double)k;
    return b > 4e18 ? (int64_t)4e18 : (int64_t)(b + 0.5);
}
// where the schedule stores the next snapshot when reversing steps [a, b) with s free
// snapshots: the right part is reversible with s-1 snapshots and the left part with s at
// one repetition less (β(s,r) = β(s-1,r) + β(s,r-1))
inline int64_t revolve_split(int64_t a, int64_t b, int s) {
    const int64_t n = b - a;
    int r = 0;
    while (binom_beta(s, r) < n) ++r;
    const int64_t left = std::max<int64_t>(1, n - binom_beta(s - 1, r));
    return a + left;
}
}  // namespace swmh

// ====================================================================== ensembles
//
// An ensemble advances several contexts together.  A step of every member is planned on
// the host (the members' kernels recorded, not launched, with their slots and scalars);
// all members run the same kernel sequence, so each kernel of the step becomes one batched
// launch whose waves find their member in a table of tile counts.  The packed plan (every
// member's arguments and the tables) lives in one of two device buffers: a context's slot
// roles repeat every two steps, so two buffers hold the steady state, and two steps without
// host work (tracer, NaN check) replay one captured hipGraph that reads them.

struct swm_ens {
    std::vector<swm_ctx *> m;
    std::string err;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;     // the seam strips of periodic pairs, beside the other strips
    hipEvent_t fork = nullptr, join = nullptr;
    int fmt = SWM_F32;
    bool per = false;
    struct Plan {                      // one step of every member, packed for upload
        std::vector<int> kind, ma, mb; // per kernel of the step
        std::vector<size_t> args_off, tab_off, qtab_off;
        std::vector<int> tiles, qblocks;
        std::vector<char> blob;        // what the device buffer holds
        std::vector<int> next_pu, next_pe, next_uv;   // member slots after the step
        std::vector<int> init;         // members whose kept q, KEu+KEv the step forms from their new state first
    };
    struct Buf {
        Plan plan;
        void *dev = nullptr;
        size_t cap = 0;
        bool valid = false;
    };
    Buf buf[2];
    int last = 1;                      // buffer of the previous eager step
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool graph_ok = false;
    int persistent = 0;                // run the steps in one persistent launch where the form allows (2: and time its phases)
    double phase_ms[4] = {-1, -1, -1, -1};   // its phases' average durations (first pair, second, dissipation, step)
    void *pdev = nullptr;              // its member records, control word and NaN flags
    size_t pcap = 0;
    int64_t persist_steps = 0;         // steps of the last swm_ens_step that ran in it
    const char *persist_why = "";      // why the last swm_ens_step did not use it

    int fail(int code, const char *fmt_, ...) {
        char b[512];
        va_list ap;
        va_start(ap, fmt_);
        vsnprintf(b, sizeof b, fmt_, ap);
        va_end(ap);
        err = b;
        return code;
    }
};


// Entry points instantiated for one number-format bundle (swm_launch.h), one table per
// bundle, each built in its own translation unit (swm_f64.hip, swm_f32.hip, swm_f16.hip,
// swm_bf16.hip); the C ABI dispatches on the context's format.  The adjoint's entries are
// null for the bundles it does not support.
struct FmtOps {
    void (*segment_lengths)(swm_ctx *ctx);
    int (*scale_static)(swm_ctx *ctx);                       // Fx, Fy, f_q ×s (velocities carried scaled)
    int (*upload_fix)(swm_ctx *ctx, int pu, int pe);         // scale, clear compensations, ghost points
    int (*run_steps)(swm_ctx *ctx, int64_t nsteps, int64_t i0, int64_t t0, int64_t *done, int64_t *nan_step);
    int (*stages)(swm_ctx *ctx, int32_t n_stages, int64_t t0);
    int (*time_kernel)(swm_ctx *ctx, int32_t which, int32_t reps, double *ms);
    int (*reduce_launch)(swm_ctx *ctx, int pu, int pe, double *part);   // row partials (mass, energy) into part
    int (*ghost)(swm_ctx *ctx);                              // ghost_points! of Prog and sst in place
    int (*dissipation)(swm_ctx *ctx);                        // Prog as (u0, v0, η0) -> dissipated Prog
    int (*tracer)(swm_ctx *ctx, int64_t i);                  // tracer! at step i with Prog's velocities
    int (*ens_tune)(swm_ens *E, int32_t reps, int *Ls, int *Lf, int *Ld);
    int (*ens_step)(swm_ens *E, int64_t nsteps, int64_t i0, int64_t *done, int64_t *nan_steps, int64_t *gsteps);
    int (*adj_forward)(swm_ctx *ctx, int64_t nsteps, int64_t i0, int64_t t0, int32_t snapshots);
    int (*adj_pullback)(swm_ctx *ctx, const void *yu, const void *yv, const void *ye, const void *ys, void *gu, void *gv,
                        void *ge, void *gs);
    int (*adj_time)(swm_ctx *ctx, int32_t which, int32_t reps, double *ms);
    int (*adj_download_fx)(swm_ctx *ctx, void *xbar_Fx);
    int (*setup_static)(swm_ctx *ctx, const swm_setup *s);   // the static fields computed on the device
};
const FmtOps &ops_f64();
const FmtOps &ops_f32();
const FmtOps &ops_f16();
const FmtOps &ops_bf16();
const FmtOps &ops_f16a();
const FmtOps &ops_bf16a();
// records an event on the context's stream when constructed and when destroyed, into kind k
// of the context's step trace (nothing when no trace runs or an ensemble records a plan)
struct KTrace {
    swm_ctx *ctx;
    int k;
    static void mark(swm_ctx *c, int kind) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return;
        (void)hipEventRecord(e, c->stream);
        c->strace->ev[kind].push_back(e);
    }
    KTrace(swm_ctx *c, int kind) : ctx(c->strace && !c->rec ? c : nullptr), k(kind) {
        if (ctx) mark(ctx, k);
    }
    ~KTrace() {
        if (ctx) mark(ctx, k);
    }
    KTrace(const KTrace &) = delete;
    KTrace &operator=(const KTrace &) = delete;
};

inline const FmtOps &fmt_ops(int fmt) {
    switch (fmt) {
    case SWM_F64: return ops_f64();
    case SWM_F32: return ops_f32();
    case SWM_F16: return ops_f16();
    case FMT_F16A: return ops_f16a();
    case FMT_BF16A: return ops_bf16a();
    default: return ops_bf16();
    }
}
