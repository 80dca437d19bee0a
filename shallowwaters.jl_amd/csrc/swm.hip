// swm.hip -- C ABI (include/swm.h) and step driver of the gfx950 ShallowWaters time step.
//
// The driver replays the loop of time_integration.jl:45-243 as kernel launches on one HIP
// stream.  The reference's arrays u, u0, u1 (and v, η alike) become four device slots per
// field whose roles rotate instead of being copied (copyto! at :49-51 and :240-242):
//   role 0  Prog        state at the start of the step
//   role 1  u0          RK accumulator (axb!), then dissipation input
//   role 2/3  u1        stage inputs, ping-pong (a stencil cannot update in place)
// Dissipation writes the new Prog into a free slot, so the previous Prog stays intact until
// the step is accepted: when NaNs are detected the loop stops before the copy-back exactly
// as the reference does (time_integration.jl:235-237).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/swm.h"
#include "swm_device.h"

using namespace swm;

#define HIPCHK(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) return ctx->fail(SWM_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

namespace {

struct Field {            // one padded device array
    void *base = nullptr;  // allocation
    char *p0 = nullptr;    // element (0,0)
};

}  // namespace

struct swm_ctx {
    swm_config c;
    std::string err;
    int dev = 0;
    hipStream_t stream = nullptr;
    int es = 4;            // element size of the compute format
    int P = 0, rows = 0;   // pitch (elements) and allocated rows of 2-D fields
    size_t elems = 0;
    int ep = 0, nux = 0, nuy = 0, nvx = 0, nvy = 0;
    int L = 64;            // rows per wave segment
    Field u[4], v[4], e[4], sst[2], um, vm, H, Fx, Fy, fq, eta_ref, Feta, qend;
    int Fx_sx = 1, Fx_sy = 0, Fy_sx = 1, Fy_sy = 0, fq_sx = 1, fq_sy = 0;
    int ruv[4] = {0, 1, 2, 3};   // slot of role k for u and v
    int re[4] = {0, 1, 2, 3};    // slot of role k for η
    int rsst = 0;
    int rk_u0 = 1, rk_u1 = 2, rk_e0 = 1, rk_e1 = 2;  // slots of u0/u1 after swm_stages
    int *dflags = nullptr;       // [0] NaN flag, [1] tracer index error
    double *dpart = nullptr;
    bool have_static = false, have_prog = false;
    std::vector<void *> allocs;

    int fail(int code, const char *fmt, ...) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        err = buf;
        return code;
    }
    template <class R> R *at(const Field &f) const { return (R *)f.p0; }
};

namespace {

int alloc_field(swm_ctx *ctx, Field &f, size_t bytes_elems, bool two_d) {
    const size_t n = two_d ? ctx->elems : bytes_elems;
    void *p = nullptr;
    HIPCHK(hipMalloc(&p, n * ctx->es));
    HIPCHK(hipMemsetAsync(p, 0, n * ctx->es, ctx->stream));
    ctx->allocs.push_back(p);
    f.base = p;
    f.p0 = (char *)p + (two_d ? ((size_t)OX + (size_t)OY * ctx->P) : (size_t)OY) * ctx->es;
    return SWM_OK;
}

// element (x, y) of a 2-D field, as a byte pointer
inline char *elem(const swm_ctx *ctx, const Field &f, int x, int y) {
    return f.p0 + ((ptrdiff_t)x + (ptrdiff_t)y * ctx->P) * ctx->es;
}

// copy a Julia (m, n) host array whose element [1,1] sits at (x0, y0)
int put2d(swm_ctx *ctx, const Field &f, const void *host, int m, int n, int x0, int y0) {
    HIPCHK(hipMemcpy2DAsync(elem(ctx, f, x0, y0), (size_t)ctx->P * ctx->es, host, (size_t)m * ctx->es,
                            (size_t)m * ctx->es, n, hipMemcpyHostToDevice, ctx->stream));
    return SWM_OK;
}
int get2d(swm_ctx *ctx, const Field &f, void *host, int m, int n, int x0, int y0) {
    HIPCHK(hipMemcpy2DAsync(host, (size_t)m * ctx->es, elem(ctx, f, x0, y0), (size_t)ctx->P * ctx->es,
                            (size_t)m * ctx->es, n, hipMemcpyDeviceToHost, ctx->stream));
    return SWM_OK;
}

// A static field that does not vary along x is kept as one value per row (stride 0 in x).
int put_static(swm_ctx *ctx, Field &f, int *sx, int *sy, const void *host, int m, int n, int x0, int y0) {
    const char *h = (const char *)host;
    bool xinv = true;
    for (int j = 0; j < n && xinv; ++j)
        for (int i = 1; i < m && xinv; ++i)
            xinv = std::memcmp(h + ((size_t)j * m + i) * ctx->es, h + (size_t)j * m * ctx->es, ctx->es) == 0;
    if (xinv) {
        int rc = alloc_field(ctx, f, (size_t)ctx->c.ny + 2 * OY + 16, false);
        if (rc) return rc;
        std::vector<char> col((size_t)n * ctx->es);
        for (int j = 0; j < n; ++j) std::memcpy(&col[(size_t)j * ctx->es], h + (size_t)j * m * ctx->es, ctx->es);
        HIPCHK(hipMemcpyAsync(f.p0 + (ptrdiff_t)y0 * ctx->es, col.data(), col.size(), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        *sx = 0;
        *sy = 1;
        return SWM_OK;
    }
    int rc = alloc_field(ctx, f, 0, true);
    if (rc) return rc;
    *sx = 1;
    *sy = ctx->P;
    return put2d(ctx, f, host, m, n, x0, y0);
}

inline int nblocks_waves(int nwaves) { return (nwaves + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK; }

template <class R> struct Launch {
    static int nstrips(const swm_ctx *ctx) { return (ctx->c.nx + 1 + WOUT - 1) / WOUT; }
    static int nseg(const swm_ctx *ctx) { return (ctx->c.ny + ctx->L - 1) / ctx->L; }

    static Static<R> statics(const swm_ctx *ctx) {
        Static<R> s;
        s.H = ctx->at<R>(ctx->H);
        s.Fx = ctx->at<R>(ctx->Fx); s.Fx_sx = ctx->Fx_sx; s.Fx_sy = ctx->Fx_sy;
        s.Fy = ctx->at<R>(ctx->Fy); s.Fy_sx = ctx->Fy_sx; s.Fy_sy = ctx->Fy_sy;
        s.fq = ctx->at<R>(ctx->fq); s.fq_sx = ctx->fq_sx; s.fq_sy = ctx->fq_sy;
        return s;
    }

    // stage `k` (1-based) of the classic RK scheme (time_integration.jl:55-78)
    static int rk_stage(swm_ctx *ctx, int k, int in_u, int in_e, int out_u, int out_e, R Fxt, R Fyt) {
        const swm_config &c = ctx->c;
        const bool per = c.bc == SWM_PERIODIC, first = k == 1, last = k == c.RKo;
        StageArgs<R> A;
        A.in_u = ctx->at<R>(ctx->u[in_u]); A.in_v = ctx->at<R>(ctx->v[in_u]); A.in_e = ctx->at<R>(ctx->e[in_e]);
        A.pr_u = ctx->at<R>(ctx->u[ctx->ruv[0]]); A.pr_v = ctx->at<R>(ctx->v[ctx->ruv[0]]);
        A.pr_e = ctx->at<R>(ctx->e[ctx->re[0]]);
        A.acc_u = ctx->at<R>(ctx->u[ctx->ruv[1]]); A.acc_v = ctx->at<R>(ctx->v[ctx->ruv[1]]);
        A.acc_e = ctx->at<R>(ctx->e[ctx->re[1]]);
        A.out_u = last ? nullptr : ctx->at<R>(ctx->u[out_u]);
        A.out_v = last ? nullptr : ctx->at<R>(ctx->v[out_u]);
        A.out_e = last ? nullptr : ctx->at<R>(ctx->e[out_e]);
        A.qend = ctx->at<R>(ctx->qend);
        A.st = statics(ctx);
        A.a = (R)c.RKaDt[k - 1];
        A.b = last ? R(0) : (R)c.RKbDt[k - 1];
        A.Fxt = Fxt; A.Fyt = Fyt; A.g = (R)c.g; A.oma = (R)c.one_minus_alpha;
        A.nx = c.nx; A.ny = c.ny; A.P = ctx->P; A.L = ctx->L;
        A.nstrips = nstrips(ctx); A.nseg = nseg(ctx);
        if (per) {
            qend_kernel<R><<<(c.ny + 1 + 255) / 256, 256, 0, ctx->stream>>>(
                A.in_u, A.in_v, A.in_e, A.st.H, A.st.fq, A.st.fq_sx, A.st.fq_sy, ctx->at<R>(ctx->qend), c.nx, c.ny, ctx->P);
        }
        const dim3 grid(nblocks_waves(A.nstrips * A.nseg)), block(WAVE * WAVES_PER_BLOCK);
#define SWM_STAGE(PER_, F_, L_) hipLaunchKernelGGL((rk_stage_kernel<R, PER_, F_, L_>), grid, block, 0, ctx->stream, A)
        if (per) {
            if (first && last) SWM_STAGE(true, true, true);
            else if (first) SWM_STAGE(true, true, false);
            else if (last) SWM_STAGE(true, false, true);
            else SWM_STAGE(true, false, false);
        } else {
            if (first && last) SWM_STAGE(false, true, true);
            else if (first) SWM_STAGE(false, true, false);
            else if (last) SWM_STAGE(false, false, true);
            else SWM_STAGE(false, false, false);
        }
#undef SWM_STAGE
        HIPCHK(hipGetLastError());
        return SWM_OK;
    }

    static int dissipation(swm_ctx *ctx, int in_uv, int in_e, int out_uv) {
        const swm_config &c = ctx->c;
        DissArgs<R> A;
        A.u0 = ctx->at<R>(ctx->u[in_uv]); A.v0 = ctx->at<R>(ctx->v[in_uv]); A.e0 = ctx->at<R>(ctx->e[in_e]);
        A.out_u = ctx->at<R>(ctx->u[out_uv]); A.out_v = ctx->at<R>(ctx->v[out_uv]);
        A.H = ctx->at<R>(ctx->H);
        A.cD = (R)c.cD; A.rD = (R)c.rD; A.cS = (R)c.cSmag; A.nuB = (R)c.nuB; A.dtd = (R)c.Dt_diff;
        A.oma = (R)c.one_minus_alpha;
        A.nx = c.nx; A.ny = c.ny; A.P = ctx->P; A.L = ctx->L;
        A.nstrips = nstrips(ctx); A.nseg = nseg(ctx);
        const dim3 grid(nblocks_waves(A.nstrips * A.nseg)), block(WAVE * WAVES_PER_BLOCK);
        const bool per = c.bc == SWM_PERIODIC, smag = c.diffusion == SWM_DIFF_SMAGORINSKY;
#define SWM_DISS(PER_, D_, S_) hipLaunchKernelGGL((dissipation_kernel<R, PER_, D_, S_>), grid, block, 0, ctx->stream, A)
#define SWM_DISS_D(PER_, S_)                      \
    switch (c.bottom_drag) {                      \
    case SWM_DRAG_QUADRATIC: SWM_DISS(PER_, 0, S_); break; \
    case SWM_DRAG_LINEAR: SWM_DISS(PER_, 1, S_); break;    \
    default: SWM_DISS(PER_, 2, S_); break;                 \
    }
        if (per) {
            if (smag) { SWM_DISS_D(true, true) } else { SWM_DISS_D(true, false) }
        } else {
            if (smag) { SWM_DISS_D(false, true) } else { SWM_DISS_D(false, false) }
        }
#undef SWM_DISS_D
#undef SWM_DISS
        HIPCHK(hipGetLastError());
        return SWM_OK;
    }

    static int tracer(swm_ctx *ctx, int64_t i, int uv_slot) {
        const swm_config &c = ctx->c;
        if (!c.tracer_advection) return SWM_OK;
        const size_t n = ctx->elems;
        if ((i + c.nadvstep_half) % c.nadvstep == 0) {      // tracer_advection.jl:12-17
            copy_kernel<R><<<1024, 256, 0, ctx->stream>>>((R *)ctx->um.base, (const R *)ctx->u[uv_slot].base, n);
            copy_kernel<R><<<1024, 256, 0, ctx->stream>>>((R *)ctx->vm.base, (const R *)ctx->v[uv_slot].base, n);
        }
        if (i % c.nadvstep == 0) {                          // tracer_advection.jl:19-42
            const int src = ctx->rsst, dst = 1 - ctx->rsst;
            int rc = tracer_advect(ctx, uv_slot, src, dst);
            if (rc) return rc;
            ctx->rsst = dst;
        }
        return SWM_OK;
    }

    static int tracer_advect(swm_ctx *ctx, int uv_slot, int src, int dst) {
        const swm_config &c = ctx->c;
        const dim3 grid((c.nx + 127) / 128, c.ny), block(128);
        const bool per = c.bc == SWM_PERIODIC;
        auto k = per ? tracer_kernel<R, true> : tracer_kernel<R, false>;
        hipLaunchKernelGGL(k, grid, block, 0, ctx->stream, ctx->at<R>(ctx->u[uv_slot]), ctx->at<R>(ctx->v[uv_slot]),
                           ctx->at<R>(ctx->um), ctx->at<R>(ctx->vm), ctx->at<R>(ctx->sst[src]), ctx->at<R>(ctx->sst[dst]),
                           (R)c.half_dtadvu, (R)c.half_dtadvv, (R)c.dtadvu, (R)c.dtadvv, c.nx, c.ny, ctx->P, ctx->ep,
                           c.tracer_consumption, (R)c.jSST, (R)c.SSTmin, ctx->dflags + 1);
        HIPCHK(hipGetLastError());
        return SWM_OK;
    }

    static int time_kernel(swm_ctx *ctx, int which, int reps, double *ms) {
        const swm_config &c = ctx->c;
        const int *ru = ctx->ruv, *re = ctx->re;
        auto launch = [&]() -> int {
            switch (which) {
            case 0: return c.RKo >= 3 ? rk_stage(ctx, 2, ru[2], re[2], ru[3], re[3], R(1), R(1))
                                      : rk_stage(ctx, 1, ru[0], re[0], ru[2], re[2], R(1), R(1));
            case 1: return rk_stage(ctx, 1, ru[0], re[0], ru[2], re[2], R(1), R(1));
            case 2: return rk_stage(ctx, c.RKo, ru[2], re[2], ru[3], re[3], R(1), R(1));
            case 3: return dissipation(ctx, ru[1], re[1], ru[3]);
            case 4: return tracer_advect(ctx, ru[0], ctx->rsst, 1 - ctx->rsst);
            }
            return ctx->fail(SWM_E_ARG, "unknown kernel %d", which);
        };
        int rc = launch();                                   // warm-up
        if (rc) return rc;
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, ctx->stream));
        for (int k = 0; k < reps && rc == SWM_OK; ++k) rc = launch();
        HIPCHK(hipEventRecord(e1, ctx->stream));
        HIPCHK(hipEventSynchronize(e1));
        float t = 0.0f;
        HIPCHK(hipEventElapsedTime(&t, e0, e1));
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        *ms = (double)t / reps;
        return rc;
    }

    static int ghosts(swm_ctx *ctx, int uv_slot, int e_slot, int sst_slot) {
        const swm_config &c = ctx->c;
        const bool per = c.bc == SWM_PERIODIC;
        auto k = per ? ghost_kernel<R, true> : ghost_kernel<R, false>;
        R *u = ctx->at<R>(ctx->u[uv_slot]), *v = ctx->at<R>(ctx->v[uv_slot]), *e = ctx->at<R>(ctx->e[e_slot]);
        R *s = ctx->at<R>(ctx->sst[sst_slot]);
        hipLaunchKernelGGL(k, dim3((c.ny + 4 + 127) / 128), dim3(128), 0, ctx->stream, u, v, e, s, c.nx, c.ny, ctx->P,
                           (R)c.one_minus_alpha, 0);
        hipLaunchKernelGGL(k, dim3((c.nx + 5 + 127) / 128), dim3(128), 0, ctx->stream, u, v, e, s, c.nx, c.ny, ctx->P,
                           (R)c.one_minus_alpha, 1);
        HIPCHK(hipGetLastError());
        return SWM_OK;
    }

    // NaN check of Prog (u, v, η of slot uv/e) and sst, feedback.jl:61-72
    static int nan_check(swm_ctx *ctx, int uv_slot, int e_slot, int sst_slot, bool *found) {
        const swm_config &c = ctx->c;
        NanArgs A;
        A.f[0] = ctx->at<R>(ctx->u[uv_slot]); A.x0[0] = -1 - ctx->ep; A.x1[0] = ctx->nux + 2 - ctx->ep; A.y0[0] = -1; A.y1[0] = c.ny + 2;
        A.f[1] = ctx->at<R>(ctx->v[uv_slot]); A.x0[1] = -1; A.x1[1] = c.nx + 2; A.y0[1] = -1; A.y1[1] = c.ny + 1;
        A.f[2] = ctx->at<R>(ctx->e[e_slot]); A.x0[2] = 0; A.x1[2] = c.nx + 1; A.y0[2] = 0; A.y1[2] = c.ny + 1;
        A.f[3] = ctx->at<R>(ctx->sst[sst_slot]); A.x0[3] = 0; A.x1[3] = c.nx + 1; A.y0[3] = 1; A.y1[3] = c.ny;
        A.P = ctx->P;
        A.nf = 4;
        HIPCHK(hipMemsetAsync(ctx->dflags, 0, sizeof(int), ctx->stream));
        nan_kernel<R><<<dim3(1, c.ny + 4, 4), 256, 0, ctx->stream>>>(A, ctx->dflags);
        HIPCHK(hipGetLastError());
        int flag = 0;
        HIPCHK(hipMemcpyAsync(&flag, ctx->dflags, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        *found = flag != 0;
        return SWM_OK;
    }

    // Stages of one RK step; leaves u0 in role 1 and the last stage input in rk_u1.
    static int rk_stages(swm_ctx *ctx, int n_stages, int64_t t) {
        const swm_config &c = ctx->c;
        const R Fxt = c.seasonal_wind_x ? (R)std::sin(c.omegaFx * (double)t) : R(1);
        const R Fyt = c.seasonal_wind_y ? (R)std::sin(c.omegaFy * (double)t) : R(1);
        int cur_u = ctx->ruv[0], cur_e = ctx->re[0];
        int nxt_u = ctx->ruv[2], nxt_e = ctx->re[2];
        for (int k = 1; k <= n_stages; ++k) {
            int rc = rk_stage(ctx, k, cur_u, cur_e, nxt_u, nxt_e, Fxt, Fyt);
            if (rc) return rc;
            if (k < c.RKo) {
                const int pu = cur_u, pe = cur_e;
                cur_u = nxt_u;
                cur_e = nxt_e;
                nxt_u = (k == 1) ? ctx->ruv[3] : pu;
                nxt_e = (k == 1) ? ctx->re[3] : pe;
            }
        }
        ctx->rk_u0 = ctx->ruv[1];
        ctx->rk_e0 = ctx->re[1];
        ctx->rk_u1 = cur_u;
        ctx->rk_e1 = cur_e;
        return SWM_OK;
    }

    static int step(swm_ctx *ctx, int64_t i, int64_t t, bool *nan) {
        const swm_config &c = ctx->c;
        *nan = false;
        int rc = rk_stages(ctx, c.RKo, t);
        if (rc) return rc;
        const int p_uv = ctx->ruv[0], a_uv = ctx->ruv[1], p_e = ctx->re[0], a_e = ctx->re[1];
        // a free u/v slot for the dissipation output: the stage slot not holding u1
        const int free_uv = (ctx->rk_u1 == ctx->ruv[2]) ? ctx->ruv[3] : ctx->ruv[2];
        int new_uv = a_uv;
        if (i % c.nstep_diff == 0) {                         // time_integration.jl:216-221
            rc = dissipation(ctx, a_uv, a_e, free_uv);
            if (rc) return rc;
            new_uv = free_uv;
        }
        const int old_sst = ctx->rsst;
        rc = tracer(ctx, i, new_uv);                         // :226-228
        if (rc) return rc;
        if (c.nout > 0 && i % c.nout == 0) {                 // feedback.jl:136-148
            rc = nan_check(ctx, p_uv, p_e, ctx->rsst, nan);
            if (rc) return rc;
            if (*nan) return SWM_OK;                         // keep Prog: no copy-back
        }
        (void)old_sst;
        // copy-back by role rotation (:240-242)
        int others_uv[3], k = 0;
        for (int s = 0; s < 4; ++s)
            if (ctx->ruv[s] != new_uv) others_uv[k++] = ctx->ruv[s];
        ctx->ruv[0] = new_uv;
        ctx->ruv[1] = others_uv[0];
        ctx->ruv[2] = others_uv[1];
        ctx->ruv[3] = others_uv[2];
        int others_e[3];
        k = 0;
        for (int s = 0; s < 4; ++s)
            if (ctx->re[s] != a_e) others_e[k++] = ctx->re[s];
        ctx->re[0] = a_e;
        ctx->re[1] = others_e[0];
        ctx->re[2] = others_e[1];
        ctx->re[3] = others_e[2];
        (void)p_uv;
        return SWM_OK;
    }
};

int check_config(swm_ctx *ctx) {
    const swm_config &c = ctx->c;
    if (c.abi_version != SWM_ABI_VERSION) return ctx->fail(SWM_E_ARG, "abi_version %d != %d", c.abi_version, SWM_ABI_VERSION);
    if (c.nx < 3 || c.ny < 3) return ctx->fail(SWM_E_ARG, "grid %dx%d too small (need nx, ny >= 3)", c.nx, c.ny);
    if (c.bc != SWM_PERIODIC && c.bc != SWM_NONPERIODIC) return ctx->fail(SWM_E_ARG, "bc %d", c.bc);
    if (!(c.T == c.Tprog && c.Tprog == c.Tcomm) || (c.T != SWM_F64 && c.T != SWM_F32))
        return ctx->fail(SWM_E_UNSUPPORTED, "formats T=%d Tprog=%d Tcomm=%d: this build supports T=Tprog=Tcomm in {Float64, Float32}",
                         c.T, c.Tprog, c.Tcomm);
    if (c.time_scheme != SWM_RK) return ctx->fail(SWM_E_UNSUPPORTED, "time_scheme %d: only RK is implemented", c.time_scheme);
    if (c.RKo < 2 || c.RKo > 4) return ctx->fail(SWM_E_ARG, "RKo %d", c.RKo);
    if (c.nstep_advcor != 0) return ctx->fail(SWM_E_UNSUPPORTED, "nstep_advcor > 0 is not implemented");
    if (c.adv_scheme != SWM_ADV_ARAKAWAHSU) return ctx->fail(SWM_E_UNSUPPORTED, "adv_scheme Sadourny is not implemented");
    if (c.dynamics != SWM_NONLINEAR) return ctx->fail(SWM_E_UNSUPPORTED, "linear dynamics is not implemented");
    if (c.continuity != SWM_CONT_ITSELF) return ctx->fail(SWM_E_UNSUPPORTED, "surface relaxation/forcing is not implemented");
    if (c.nstep_diff < 1 || (c.tracer_advection && c.nadvstep < 1)) return ctx->fail(SWM_E_ARG, "nstep_diff/nadvstep must be >= 1");
    if (c.bottom_drag < 0 || c.bottom_drag > 2 || c.diffusion < 0 || c.diffusion > 1)
        return ctx->fail(SWM_E_ARG, "bottom_drag %d / diffusion %d", c.bottom_drag, c.diffusion);
    return SWM_OK;
}

template <class R> int run_steps(swm_ctx *ctx, int64_t nsteps, int64_t i0, int64_t t0, int64_t *done, int64_t *nan_step) {
    *done = 0;
    *nan_step = 0;
    for (int64_t k = 0; k < nsteps; ++k) {
        bool nan = false;
        int rc = Launch<R>::step(ctx, i0 + k, t0 + k * ctx->c.dtint, &nan);
        if (rc) return rc;
        if (nan) {
            *nan_step = i0 + k;
            break;
        }
        *done = k + 1;
    }
    int flags[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(flags, ctx->dflags, sizeof flags, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (flags[1]) {
        int z = 0;
        HIPCHK(hipMemcpy(ctx->dflags + 1, &z, sizeof z, hipMemcpyHostToDevice));
        return ctx->fail(SWM_E_ARG, "tracer departure point is NaN or outside the grid (InexactError/BoundsError in the reference)");
    }
    return SWM_OK;
}

}  // namespace

// ====================================================================== C ABI

extern "C" {

int swm_version(int32_t *abi_version) {
    if (!abi_version) return SWM_E_ARG;
    *abi_version = SWM_ABI_VERSION;
    return SWM_OK;
}

int swm_device_count(int32_t *ndev) {
    if (!ndev) return SWM_E_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *ndev = n;
    return SWM_OK;
}

const char *swm_last_error(const swm_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int swm_create(const swm_config *cfg, swm_ctx **out) {
    if (!cfg || !out) return SWM_E_ARG;
    *out = nullptr;
    swm_ctx *ctx = new swm_ctx();
    ctx->c = *cfg;
    int rc = check_config(ctx);
    if (rc) {
        fprintf(stderr, "swm_create: %s\n", ctx->err.c_str());
        delete ctx;
        return rc;
    }
    const swm_config &c = ctx->c;
    ctx->es = c.T == SWM_F64 ? 8 : 4;
    ctx->ep = c.bc == SWM_PERIODIC ? 1 : 0;
    ctx->nux = ctx->ep ? c.nx : c.nx - 1;
    ctx->nuy = c.ny;
    ctx->nvx = c.nx;
    ctx->nvy = c.ny - 1;
    ctx->P = ((c.nx + 2 * OX + 63) / 64) * 64;
    ctx->rows = c.ny + 2 * OY + 4;
    ctx->elems = (size_t)ctx->P * ctx->rows;
    ctx->L = c.ny >= 2048 ? 64 : 32;
    if (const char *s = getenv("SWM_ROWS_PER_WAVE")) {   // tuning knob: rows each wavefront marches
        const int l = atoi(s);
        if (l >= 4 && l <= 4096) ctx->L = l;
    }
    if (hipGetDevice(&ctx->dev) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        fprintf(stderr, "swm_create: no HIP device/stream\n");
        delete ctx;
        return SWM_E_HIP;
    }
    auto bail = [&](int code) {
        swm_destroy(ctx);
        return code;
    };
    for (int s = 0; s < 4; ++s) {
        if ((rc = alloc_field(ctx, ctx->u[s], 0, true)) || (rc = alloc_field(ctx, ctx->v[s], 0, true)) ||
            (rc = alloc_field(ctx, ctx->e[s], 0, true)))
            return bail(rc);
    }
    if ((rc = alloc_field(ctx, ctx->sst[0], 0, true)) || (rc = alloc_field(ctx, ctx->sst[1], 0, true)) ||
        (rc = alloc_field(ctx, ctx->um, 0, true)) || (rc = alloc_field(ctx, ctx->vm, 0, true)) ||
        (rc = alloc_field(ctx, ctx->H, 0, true)) || (rc = alloc_field(ctx, ctx->qend, (size_t)c.ny + 2 * OY + 16, false)))
        return bail(rc);
    void *p = nullptr;
    if (hipMalloc(&p, 4 * sizeof(int)) != hipSuccess) return bail(SWM_E_HIP);
    ctx->allocs.push_back(p);
    ctx->dflags = (int *)p;
    if (hipMemset(p, 0, 4 * sizeof(int)) != hipSuccess) return bail(SWM_E_HIP);
    if (hipMalloc(&p, 2 * sizeof(double) * (size_t)(c.ny + 8)) != hipSuccess) return bail(SWM_E_HIP);
    ctx->allocs.push_back(p);
    ctx->dpart = (double *)p;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return bail(SWM_E_HIP);
    *out = ctx;
    return SWM_OK;
}

int swm_destroy(swm_ctx *ctx) {
    if (!ctx) return SWM_OK;
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    for (void *p : ctx->allocs) hipFree(p);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
    return SWM_OK;
}

int swm_set_static(swm_ctx *ctx, const void *H, const void *Fx, const void *Fy, const void *f_q, const void *eta_ref,
                   const void *Feta) {
    if (!ctx) return SWM_E_ARG;
    if (!H || !Fx || !Fy || !f_q) return ctx->fail(SWM_E_ARG, "H, Fx, Fy and f_q are required");
    (void)eta_ref;
    (void)Feta;
    if (ctx->have_static) return ctx->fail(SWM_E_STATE, "static fields already set");
    const swm_config &c = ctx->c;
    int rc;
    if ((rc = put2d(ctx, ctx->H, H, c.nx + 2, c.ny + 2, 0, 0))) return rc;
    if ((rc = put_static(ctx, ctx->Fx, &ctx->Fx_sx, &ctx->Fx_sy, Fx, ctx->nux, ctx->nuy, 1 - ctx->ep, 1))) return rc;
    if ((rc = put_static(ctx, ctx->Fy, &ctx->Fy_sx, &ctx->Fy_sy, Fy, ctx->nvx, ctx->nvy, 1, 1))) return rc;
    if ((rc = put_static(ctx, ctx->fq, &ctx->fq_sx, &ctx->fq_sy, f_q, c.nx + 1, c.ny + 1, 0, 0))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->have_static = true;
    return SWM_OK;
}

// halo cells the reference never writes must hold zero (add_halo, ghost_points.jl:11-22)
static int check_never_written(swm_ctx *ctx, const void *u, const void *v) {
    const swm_config &c = ctx->c;
    const int es = ctx->es;
    auto nz = [&](const void *a, int m, int i, int j) {  // 1-based Julia index
        const char *p = (const char *)a + ((size_t)(i - 1) + (size_t)(j - 1) * m) * es;
        if (es == 8) return *(const double *)p != 0.0;
        return *(const float *)p != 0.0f;
    };
    const int mu = ctx->nux + 4, nu = ctx->nuy + 4, mv = ctx->nvx + 4, nv = ctx->nvy + 4;
    if (c.bc == SWM_NONPERIODIC) {
        for (int j = 3; j <= nu - 2; ++j)
            for (int i : {1, 2, mu - 1, mu})
                if (nz(u, mu, i, j)) return ctx->fail(SWM_E_ARG, "u[%d,%d] lies in the normal-velocity halo and must be 0", i, j);
    } else {
        for (int j : {1, 2, nu - 1, nu})
            for (int i = 3; i <= mu - 2; ++i)
                if (nz(u, mu, i, j)) return ctx->fail(SWM_E_ARG, "u[%d,%d] lies in the periodic u y-halo and must be 0", i, j);
    }
    for (int j : {1, 2, nv - 1, nv})
        for (int i = 3; i <= mv - 2; ++i)
            if (nz(v, mv, i, j)) return ctx->fail(SWM_E_ARG, "v[%d,%d] lies in the normal-velocity halo and must be 0", i, j);
    return SWM_OK;
}

int swm_upload_prog(swm_ctx *ctx, const void *u, const void *v, const void *eta, const void *sst) {
    if (!ctx) return SWM_E_ARG;
    if (!u || !v || !eta || !sst) return ctx->fail(SWM_E_ARG, "null array");
    int rc = check_never_written(ctx, u, v);
    if (rc) return rc;
    const swm_config &c = ctx->c;
    const int pu = ctx->ruv[0], pe = ctx->re[0];
    if ((rc = put2d(ctx, ctx->u[pu], u, ctx->nux + 4, ctx->nuy + 4, -1 - ctx->ep, -1))) return rc;
    if ((rc = put2d(ctx, ctx->v[pu], v, ctx->nvx + 4, ctx->nvy + 4, -1, -1))) return rc;
    if ((rc = put2d(ctx, ctx->e[pe], eta, c.nx + 2, c.ny + 2, 0, 0))) return rc;
    if ((rc = put2d(ctx, ctx->sst[ctx->rsst], sst, c.nx + 2, c.ny, 0, 1))) return rc;
    rc = c.T == SWM_F64 ? Launch<double>::ghosts(ctx, pu, pe, ctx->rsst) : Launch<float>::ghosts(ctx, pu, pe, ctx->rsst);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->have_prog = true;
    return SWM_OK;
}

int swm_download_prog(swm_ctx *ctx, void *u, void *v, void *eta, void *sst) {
    if (!ctx) return SWM_E_ARG;
    if (!ctx->have_prog) return ctx->fail(SWM_E_STATE, "no prognostic state uploaded");
    const swm_config &c = ctx->c;
    int rc;
    const int pu = ctx->ruv[0], pe = ctx->re[0];
    if (u && (rc = get2d(ctx, ctx->u[pu], u, ctx->nux + 4, ctx->nuy + 4, -1 - ctx->ep, -1))) return rc;
    if (v && (rc = get2d(ctx, ctx->v[pu], v, ctx->nvx + 4, ctx->nvy + 4, -1, -1))) return rc;
    if (eta && (rc = get2d(ctx, ctx->e[pe], eta, c.nx + 2, c.ny + 2, 0, 0))) return rc;
    if (sst && (rc = get2d(ctx, ctx->sst[ctx->rsst], sst, c.nx + 2, c.ny, 0, 1))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return SWM_OK;
}

int swm_step(swm_ctx *ctx, int64_t nsteps, int64_t i0, int64_t t0, int64_t *steps_done, int64_t *nan_step) {
    if (!ctx || !steps_done || !nan_step) return SWM_E_ARG;
    if (!ctx->have_static || !ctx->have_prog) return ctx->fail(SWM_E_STATE, "set_static and upload_prog first");
    if (nsteps < 0 || i0 < 1) return ctx->fail(SWM_E_ARG, "nsteps >= 0 and i0 >= 1 required");
    return ctx->c.T == SWM_F64 ? run_steps<double>(ctx, nsteps, i0, t0, steps_done, nan_step)
                               : run_steps<float>(ctx, nsteps, i0, t0, steps_done, nan_step);
}

int swm_stages(swm_ctx *ctx, int32_t n_stages, int64_t i0, int64_t t0) {
    if (!ctx) return SWM_E_ARG;
    if (!ctx->have_static || !ctx->have_prog) return ctx->fail(SWM_E_STATE, "set_static and upload_prog first");
    if (n_stages < 1 || n_stages > ctx->c.RKo) return ctx->fail(SWM_E_ARG, "n_stages %d out of 1..%d", n_stages, ctx->c.RKo);
    (void)i0;
    int rc = ctx->c.T == SWM_F64 ? Launch<double>::rk_stages(ctx, n_stages, t0) : Launch<float>::rk_stages(ctx, n_stages, t0);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return SWM_OK;
}

int swm_download_rk(swm_ctx *ctx, void *u0, void *v0, void *eta0, void *u1, void *v1, void *eta1) {
    if (!ctx) return SWM_E_ARG;
    const swm_config &c = ctx->c;
    int rc;
    const int mu = ctx->nux + 4, nu = ctx->nuy + 4, mv = ctx->nvx + 4, nv = ctx->nvy + 4;
    if (u0 && (rc = get2d(ctx, ctx->u[ctx->rk_u0], u0, mu, nu, -1 - ctx->ep, -1))) return rc;
    if (v0 && (rc = get2d(ctx, ctx->v[ctx->rk_u0], v0, mv, nv, -1, -1))) return rc;
    if (eta0 && (rc = get2d(ctx, ctx->e[ctx->rk_e0], eta0, c.nx + 2, c.ny + 2, 0, 0))) return rc;
    if (u1 && (rc = get2d(ctx, ctx->u[ctx->rk_u1], u1, mu, nu, -1 - ctx->ep, -1))) return rc;
    if (v1 && (rc = get2d(ctx, ctx->v[ctx->rk_u1], v1, mv, nv, -1, -1))) return rc;
    if (eta1 && (rc = get2d(ctx, ctx->e[ctx->rk_e1], eta1, c.nx + 2, c.ny + 2, 0, 0))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return SWM_OK;
}

int swm_time_kernel(swm_ctx *ctx, int32_t which, int32_t reps, double *ms_avg) {
    if (!ctx || !ms_avg) return SWM_E_ARG;
    if (!ctx->have_static || !ctx->have_prog) return ctx->fail(SWM_E_STATE, "set_static and upload_prog first");
    if (reps < 1) return ctx->fail(SWM_E_ARG, "reps must be >= 1");
    return ctx->c.T == SWM_F64 ? Launch<double>::time_kernel(ctx, which, reps, ms_avg)
                               : Launch<float>::time_kernel(ctx, which, reps, ms_avg);
}

int swm_reduce(swm_ctx *ctx, double *mass, double *energy) {
    if (!ctx || !mass || !energy) return SWM_E_ARG;
    if (!ctx->have_static || !ctx->have_prog) return ctx->fail(SWM_E_STATE, "set_static and upload_prog first");
    const swm_config &c = ctx->c;
    const int pu = ctx->ruv[0], pe = ctx->re[0];
    if (c.T == SWM_F64)
        reduce_kernel<double><<<c.ny, 256, 0, ctx->stream>>>(ctx->at<double>(ctx->u[pu]), ctx->at<double>(ctx->v[pu]),
                                                             ctx->at<double>(ctx->e[pe]), ctx->at<double>(ctx->H), c.nx,
                                                             c.ny, ctx->P, ctx->ep, c.g, ctx->dpart);
    else
        reduce_kernel<float><<<c.ny, 256, 0, ctx->stream>>>(ctx->at<float>(ctx->u[pu]), ctx->at<float>(ctx->v[pu]),
                                                            ctx->at<float>(ctx->e[pe]), ctx->at<float>(ctx->H), c.nx, c.ny,
                                                            ctx->P, ctx->ep, c.g, ctx->dpart);
    HIPCHK(hipGetLastError());
    std::vector<double> part(2 * (size_t)c.ny);
    HIPCHK(hipMemcpyAsync(part.data(), ctx->dpart, part.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    double m = 0.0, e = 0.0, cm = 0.0, ce = 0.0;   // Neumaier-compensated row sums
    for (int y = 0; y < c.ny; ++y) {
        const double a = part[2 * y], b = part[2 * y + 1];
        double t = m + a;
        cm += std::fabs(m) >= std::fabs(a) ? (m - t) + a : (a - t) + m;
        m = t;
        t = e + b;
        ce += std::fabs(e) >= std::fabs(b) ? (e - t) + b : (b - t) + e;
        e = t;
    }
    *mass = m + cm;
    *energy = e + ce;
    return SWM_OK;
}

}  // extern "C"
