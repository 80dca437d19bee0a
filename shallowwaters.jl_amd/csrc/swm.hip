// swm.hip -- C ABI (include/swm.h) and step driver of the gfx950 ShallowWaters time step.
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

#include <mutex>

#include "swm_host.h"

namespace {

int check_config(swm_ctx *ctx) {
    const swm_config &c = ctx->c;
    if (c.abi_version != SWM_ABI_VERSION) return ctx->fail(SWM_E_ARG, "abi_version %d != %d", c.abi_version, SWM_ABI_VERSION);
    if (c.nx < 3 || c.ny < 3) return ctx->fail(SWM_E_ARG, "grid %dx%d too small (need nx, ny >= 3)", c.nx, c.ny);
    if (c.bc != SWM_PERIODIC && c.bc != SWM_NONPERIODIC) return ctx->fail(SWM_E_ARG, "bc %d", c.bc);
    if (c.T < SWM_F64 || c.T > SWM_BF16 || c.Tprog < SWM_F64 || c.Tprog > SWM_BF16 || c.Tcomm < SWM_F64 || c.Tcomm > SWM_BF16)
        return ctx->fail(SWM_E_ARG, "number format codes T=%d Tprog=%d Tcomm=%d", c.T, c.Tprog, c.Tcomm);
    const bool same = c.T == c.Tprog && (c.T == SWM_F64 || c.T == SWM_F32);
    const bool mixed = (c.T == SWM_F16 || c.T == SWM_BF16) && c.Tprog == SWM_F32 &&
                       (c.Tcomm == c.T || c.Tcomm == SWM_F32 || c.Tcomm == SWM_F64);
    if (!same && !mixed)
        return ctx->fail(SWM_E_UNSUPPORTED, "formats T=%d Tprog=%d Tcomm=%d: supported are T = Tprog in {Float64, Float32} "
                         "(any Tcomm), and T in {Float16, BFloat16} with Tprog = Float32 and Tcomm in {T, Tprog}",
                         c.T, c.Tprog, c.Tcomm);
    if (c.time_scheme < SWM_RK || c.time_scheme > SWM_4SSPRK3) return ctx->fail(SWM_E_ARG, "time_scheme %d", c.time_scheme);
    if (mixed && c.time_scheme != SWM_RK)
        return ctx->fail(SWM_E_UNSUPPORTED, "T != Tprog is implemented for the classic RK scheme only");
    if (c.time_scheme == SWM_RK && (c.RKo < 2 || c.RKo > 4)) return ctx->fail(SWM_E_ARG, "RKo %d", c.RKo);
    if (c.time_scheme == SWM_SSPRK2 && c.RKs < 2) return ctx->fail(SWM_E_ARG, "RKs %d", c.RKs);
    if (c.time_scheme == SWM_SSPRK3 && (c.ssp_n < 2 || c.ssp_s != c.ssp_n * c.ssp_n)) return ctx->fail(SWM_E_ARG, "RKn %d", c.ssp_n);
    if (c.nstep_advcor < 0) return ctx->fail(SWM_E_ARG, "nstep_advcor %d", c.nstep_advcor);
    if (c.scale_log2 < -20 || c.scale_log2 > 20) return ctx->fail(SWM_E_ARG, "scale_log2 %d out of -20..20", c.scale_log2);
    if (c.compensated && c.time_scheme != SWM_RK) return ctx->fail(SWM_E_UNSUPPORTED, "compensated needs the classic RK scheme");
    if (c.adv_scheme != SWM_ADV_ARAKAWAHSU && c.adv_scheme != SWM_ADV_SADOURNY) return ctx->fail(SWM_E_ARG, "adv_scheme %d", c.adv_scheme);
    if (c.dynamics != SWM_NONLINEAR && c.dynamics != SWM_LINEAR) return ctx->fail(SWM_E_ARG, "dynamics %d", c.dynamics);
    if (c.continuity < SWM_CONT_ITSELF || c.continuity > SWM_CONT_FORCING) return ctx->fail(SWM_E_ARG, "continuity %d", c.continuity);
    if (c.nstep_diff < 1 || (c.tracer_advection && c.nadvstep < 1)) return ctx->fail(SWM_E_ARG, "nstep_diff/nadvstep must be >= 1");
    if (c.bottom_drag < 0 || c.bottom_drag > 2 || c.diffusion < 0 || c.diffusion > 1)
        return ctx->fail(SWM_E_ARG, "bottom_drag %d / diffusion %d", c.bottom_drag, c.diffusion);
    if (c.t_arith32 != 0 && c.t_arith32 != 1) return ctx->fail(SWM_E_ARG, "t_arith32 %d", c.t_arith32);
    if (c.t_arith32 && !mixed)
        return ctx->fail(SWM_E_UNSUPPORTED, "t_arith32 needs T in {Float16, BFloat16} with Tprog = Float32");
    if (c.t_arith32 && c.nstep_advcor != 0)
        return ctx->fail(SWM_E_UNSUPPORTED, "t_arith32 with nstep_advcor > 0 (q and KE kept in T) is not implemented");
    ctx->mixed = mixed;
    ctx->fmt = c.t_arith32 ? (c.T == SWM_F16 ? FMT_F16A : FMT_BF16A) : c.T;
    return SWM_OK;
}

// the members' work goes to the ensemble's stream while it runs
// ---- the static scratch of swm_reduce_noalloc
//
// The row partials (mass, energy) of up to NOALLOC_MAX_ROWS rows and, after them, the two
// compensated sums, in device memory the library holds for the whole process (one copy on
// each GPU it runs on, 512 KiB).  Its device address is looked up once per device, with that
// device current, when the first context is created there; host code passes that address,
// never the array's name, to kernels and copies.  Every call on a device shares its copy.
constexpr int NOALLOC_MAX_ROWS = 32768;
constexpr int MAX_DEVICES = 64;
static __device__ __align__(256) double scratch[2 * NOALLOC_MAX_ROWS + 2];
double *g_scratch[MAX_DEVICES] = {};
std::mutex g_scratch_mu;

// scratch's address on the current device ctx->dev (set up once per device)
int scratch_setup(swm_ctx *ctx) {
    if (ctx->dev < 0 || ctx->dev >= MAX_DEVICES) return ctx->fail(SWM_E_HIP, "device %d out of range", ctx->dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[ctx->dev]) return SWM_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return ctx->fail(SWM_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", ctx->dev, hipGetErrorString(e));
    g_scratch[ctx->dev] = (double *)p;
    return SWM_OK;
}

// The rows' partials summed in row order with Neumaier's compensation, by one lane: the
// same operations in the same order as swm_reduce's host loop, so the same doubles.
__global__ void reduce_rows_kernel(const double *part, int ny, double *out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double m = 0.0, e = 0.0, cm = 0.0, ce = 0.0;
    for (int y = 0; y < ny; ++y) {
        const double a = part[2 * y], b = part[2 * y + 1];
        double t = m + a;
        cm += fabs(m) >= fabs(a) ? (m - t) + a : (a - t) + m;
        m = t;
        t = e + b;
        ce += fabs(e) >= fabs(b) ? (e - t) + b : (b - t) + e;
        e = t;
    }
    out[0] = m + cm;
    out[1] = e + ce;
}

// the adjoint's number formats: T = Tprog = Tcomm in {Float64, Float32}
int adj_ok(swm_ctx *ctx) {
    if ((ctx->fmt == SWM_F64 && ctx->c.Tprog == SWM_F64) || (ctx->fmt == SWM_F32 && ctx->c.Tprog == SWM_F32)) return SWM_OK;
    return ctx->fail(SWM_E_UNSUPPORTED, "the adjoint needs T = Tprog = Tcomm in {Float64, Float32}");
}

struct StreamSwap {
    swm_ens *E;
    std::vector<hipStream_t> saved;
    explicit StreamSwap(swm_ens *e) : E(e) {
        for (swm_ctx *c : E->m) { saved.push_back(c->stream); c->stream = E->stream; }
    }
    ~StreamSwap() {
        for (size_t k = 0; k < E->m.size(); ++k) E->m[k]->stream = saved[k];
    }
};

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
    ctx->es = c.Tprog == SWM_F64 ? 8 : 4;
    ctx->est = c.T == SWM_F64 ? 8 : c.T == SWM_F32 ? 4 : 2;
    ctx->ep = c.bc == SWM_PERIODIC ? 1 : 0;
    ctx->nux = ctx->ep ? c.nx : c.nx - 1;
    ctx->nuy = c.ny;
    ctx->nvx = c.nx;
    ctx->nvy = c.ny - 1;
    ctx->sc = std::ldexp(1.0, c.scale_log2);
    ctx->sci = std::ldexp(1.0, -c.scale_log2);
    ctx->P = ((c.nx + 2 * OX + 63) / 64) * 64;
    ctx->rows = c.ny + 2 * OY + 4;
    ctx->elems = (size_t)ctx->P * ctx->rows;
    if (hipGetDevice(&ctx->dev) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        fprintf(stderr, "swm_create: no HIP device/stream\n");
        delete ctx;
        return SWM_E_HIP;
    }
    if (scratch_setup(ctx) != SWM_OK) {
        fprintf(stderr, "swm_create: %s\n", ctx->err.c_str());
        swm_destroy(ctx);
        return SWM_E_HIP;
    }
    if (c.bc == SWM_PERIODIC &&
        (hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking) != hipSuccess ||
         hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming) != hipSuccess ||
         hipEventCreateWithFlags(&ctx->join, hipEventDisableTiming) != hipSuccess)) {
        fprintf(stderr, "swm_create: no second HIP stream\n");
        swm_destroy(ctx);
        return SWM_E_HIP;
    }
    fmt_ops(ctx->fmt).segment_lengths(ctx);
    // tuning overrides: rows each wavefront marches in the stage / dissipation kernels
    if (const char *s = getenv("SWM_ROWS_PER_WAVE")) {
        const int l = atoi(s);
        if (l >= 4 && l <= 4096) ctx->L = l;
    }
    if (const char *s = getenv("SWM_FUSE")) ctx->fuse = atoi(s) != 0;   // 0: one kernel per stage
    if (const char *s = getenv("SWM_ROWS_PER_WAVE_FUSED")) {
        const int l = atoi(s);
        if (l >= 4 && l <= 4096) ctx->Lf = l;
    }
    if (const char *s = getenv("SWM_ROWS_PER_WAVE_DISS")) {
        const int l = atoi(s);
        if (l >= 4 && l <= 4096) ctx->Ld = l;
    }
    auto bail = [&](int code) {
        swm_destroy(ctx);
        return code;
    };
    // the arena: Tprog fields 12 state slots + 2 sst; T fields um, vm, H, Fx, Fy, f_q
    // (+ η_ref, Fη; + 6 u1 slots with mixed formats) and the q column
    const int nq = 14 + (c.compensated ? 3 : 0);
    const int nt = 6 + (c.continuity != SWM_CONT_ITSELF ? 2 : 0) + (ctx->mixed ? 6 : 0) + (c.dynamics == SWM_LINEAR ? 2 : 0) +
                   ((adv_form(c) & ADV_KEEP) ? 2 : 0);
    const size_t n1d = (size_t)c.ny + 2 * OY + 16;
    ctx->arena_bytes = (size_t)nq * (ctx->elems * ctx->es + 256) + (size_t)nt * (ctx->elems * ctx->est + 256) +
                       n1d * (c.t_arith32 ? 4 : ctx->est) + 256;
    if (ctx->arena_bytes > 0xFFFF0000ull) {     // byte offsets into the arena are 32-bit
        ctx->fail(SWM_E_UNSUPPORTED, "grid %dx%d needs %zu bytes of device fields; one context addresses at most 4 GiB",
                  c.nx, c.ny, ctx->arena_bytes);
        fprintf(stderr, "swm_create: %s\n", ctx->err.c_str());
        return bail(SWM_E_UNSUPPORTED);
    }
    if (hipMalloc(&ctx->arena, ctx->arena_bytes) != hipSuccess) return bail(SWM_E_HIP);
    ctx->allocs.push_back(ctx->arena);
    if (hipMemsetAsync(ctx->arena, 0, ctx->arena_bytes, ctx->stream) != hipSuccess) return bail(SWM_E_HIP);
    const int es = ctx->es, est = ctx->est;
    for (int s = 0; s < 4; ++s) {
        if ((rc = alloc_field(ctx, ctx->u[s], 0, true, es)) || (rc = alloc_field(ctx, ctx->v[s], 0, true, es)) ||
            (rc = alloc_field(ctx, ctx->e[s], 0, true, es)))
            return bail(rc);
    }
    if (ctx->mixed)
        for (int s = 0; s < 2; ++s)
            if ((rc = alloc_field(ctx, ctx->tu[s], 0, true, est)) || (rc = alloc_field(ctx, ctx->tv[s], 0, true, est)) ||
                (rc = alloc_field(ctx, ctx->te[s], 0, true, est)))
                return bail(rc);
    if ((rc = alloc_field(ctx, ctx->sst[0], 0, true, es)) || (rc = alloc_field(ctx, ctx->sst[1], 0, true, es)) ||
        (rc = alloc_field(ctx, ctx->um, 0, true, est)) || (rc = alloc_field(ctx, ctx->vm, 0, true, est)) ||
        (rc = alloc_field(ctx, ctx->H, 0, true, est)) || (rc = alloc_field(ctx, ctx->Fx, 0, true, est)) ||
        (rc = alloc_field(ctx, ctx->Fy, 0, true, est)) || (rc = alloc_field(ctx, ctx->fq, 0, true, est)) ||
        (rc = alloc_field(ctx, ctx->qend, n1d, false, c.t_arith32 ? 4 : est)))
        return bail(rc);
    if (c.compensated && ((rc = alloc_field(ctx, ctx->cu, 0, true, es)) || (rc = alloc_field(ctx, ctx->cv, 0, true, es)) ||
                          (rc = alloc_field(ctx, ctx->ce, 0, true, es))))
        return bail(rc);
    if (c.continuity != SWM_CONT_ITSELF &&
        ((rc = alloc_field(ctx, ctx->eta_ref, 0, true, est)) || (rc = alloc_field(ctx, ctx->Feta, 0, true, est))))
        return bail(rc);
    if (c.dynamics == SWM_LINEAR &&
        ((rc = alloc_field(ctx, ctx->f_u, 0, true, est)) || (rc = alloc_field(ctx, ctx->f_v, 0, true, est))))
        return bail(rc);
    if ((adv_form(c) & ADV_KEEP) &&
        ((rc = alloc_field(ctx, ctx->qs, 0, true, est)) || (rc = alloc_field(ctx, ctx->ks, 0, true, est))))
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
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->adj) {
        ctx->adj->release();
        delete ctx->adj;
    }
    if (ctx->stream2) (void)hipStreamSynchronize(ctx->stream2);
    for (void *p : ctx->allocs) (void)hipFree(p);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
    if (ctx->fork) (void)hipEventDestroy(ctx->fork);
    if (ctx->join) (void)hipEventDestroy(ctx->join);
    delete ctx;
    return SWM_OK;
}

// The periodic fused stage pairs compute the columns beyond the seam from the columns they
// copy; that gives the reference's values where the copies are exact (Tcomm at least as wide
// as the conversion to T needs), H's ghost columns equal the columns they face across the
// seam, and f_q does not vary along x.  H: the host copy, Julia layout (nx+2, ny+2).
static void seam_check(swm_ctx *ctx, const char *h) {
    const swm_config &c = ctx->c;
    const bool exact = c.Tcomm == c.T || c.Tcomm == c.Tprog || c.Tcomm == SWM_F64 ||
                       (c.Tcomm == SWM_F32 && c.Tprog != SWM_F64);
    const size_t es = (size_t)ctx->est, m = (size_t)c.nx + 2;
    bool same = true;
    for (size_t j = 0; j < (size_t)c.ny + 2 && same; ++j)
        same = std::memcmp(h + (j * m + 0) * es, h + (j * m + c.nx) * es, es) == 0 &&
               std::memcmp(h + (j * m + c.nx + 1) * es, h + (j * m + 1) * es, es) == 0;
    ctx->seam_ok = exact && same && ctx->fq_sx == 0;
}

int swm_set_static(swm_ctx *ctx, const void *H, const void *Fx, const void *Fy, const void *f_q, const void *eta_ref,
                   const void *Feta) {
    if (!ctx) return SWM_E_ARG;
    if (!H || !Fx || !Fy || !f_q) return ctx->fail(SWM_E_ARG, "H, Fx, Fy and f_q are required");
    if (ctx->have_static) return ctx->fail(SWM_E_STATE, "static fields already set");
    const swm_config &c = ctx->c;
    int rc;
    if ((rc = put2d(ctx, ctx->H, H, c.nx + 2, c.ny + 2, 0, 0))) return rc;
    if ((rc = put_static(ctx, ctx->Fx, &ctx->Fx_sx, &ctx->Fx_sy, Fx, ctx->nux, ctx->nuy, 1 - ctx->ep, 1))) return rc;
    if ((rc = put_static(ctx, ctx->Fy, &ctx->Fy_sx, &ctx->Fy_sy, Fy, ctx->nvx, ctx->nvy, 1, 1))) return rc;
    if ((rc = put_static(ctx, ctx->fq, &ctx->fq_sx, &ctx->fq_sy, f_q, c.nx + 1, c.ny + 1, 0, 0))) return rc;
    if (c.scale_log2 != 0) {                     // the forcing and f of velocities carried scaled: ×s (exact)
        rc = fmt_ops(ctx->fmt).scale_static(ctx);
        if (rc) return rc;
    }
    if (c.bc == SWM_PERIODIC) seam_check(ctx, (const char *)H);
    if (c.continuity == SWM_CONT_RELAX) {
        if (!eta_ref) return ctx->fail(SWM_E_ARG, "eta_ref is required for surface relaxation");
        if ((rc = put2d(ctx, ctx->eta_ref, eta_ref, c.nx, c.ny, 1, 1))) return rc;
    }
    if (c.continuity == SWM_CONT_FORCING) {
        if (!Feta) return ctx->fail(SWM_E_ARG, "Feta is required for surface forcing");
        if ((rc = put2d(ctx, ctx->Feta, Feta, c.nx, c.ny, 1, 1))) return rc;
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->have_static = true;
    return SWM_OK;
}

int swm_set_static_device(swm_ctx *ctx, const swm_setup *st) {
    if (!ctx) return SWM_E_ARG;
    if (!st) return ctx->fail(SWM_E_ARG, "setup is required");
    if (ctx->have_static) return ctx->fail(SWM_E_STATE, "static fields already set");
    if (ctx->c.dynamics == SWM_LINEAR && ctx->have_fuv) return ctx->fail(SWM_E_STATE, "f_u, f_v already set");
    if (st->topography < SWM_TOPO_FLAT || st->topography > SWM_TOPO_SEAMOUNT)
        return ctx->fail(SWM_E_ARG, "unknown topography");
    if (st->wind_x < SWM_WIND_CHANNEL || st->wind_x > SWM_WIND_NONE) return ctx->fail(SWM_E_ARG, "unknown wind_x");
    // the other meridional winds are functions on the u-grid (forcing.jl:15-37), of another shape than Fy
    if (st->wind_y != SWM_WIND_CONSTANT && st->wind_y != SWM_WIND_NONE)
        return ctx->fail(SWM_E_ARG, "wind_y must be constant or none");
    if (!(st->Delta > 0) || !(st->Lx > 0) || !(st->Ly > 0) || !(st->H > 0) || !(st->rho > 0))
        return ctx->fail(SWM_E_ARG, "Delta, Lx, Ly, H and rho must be positive");
    const swm_config &c = ctx->c;
    int rc = fmt_ops(ctx->fmt).setup_static(ctx, st);
    if (rc) return rc;
    if (c.scale_log2 != 0 && (rc = fmt_ops(ctx->fmt).scale_static(ctx))) return rc;
    if (c.bc == SWM_PERIODIC) {
        std::vector<char> h((size_t)(c.nx + 2) * (c.ny + 2) * ctx->est);
        if ((rc = get2d(ctx, ctx->H, h.data(), c.nx + 2, c.ny + 2, 0, 0))) return rc;
        HIPCHK(hipStreamSynchronize(ctx->stream));
        seam_check(ctx, h.data());
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->have_static = true;
    if (c.dynamics == SWM_LINEAR) ctx->have_fuv = true;
    return SWM_OK;
}

int swm_download_static(swm_ctx *ctx, void *H, void *Fx, void *Fy, void *f_q, void *eta_ref, void *Feta) {
    if (!ctx) return SWM_E_ARG;
    if (!ctx->have_static) return ctx->fail(SWM_E_STATE, "static fields not set");
    const swm_config &c = ctx->c;
    int rc;
    // a field kept one value per row comes back spread along x, as it was given
    auto get = [&](const Field &f, int sx, void *host, int m, int n, int x0, int y0) -> int {
        if (!host) return SWM_OK;
        if (sx != 0) return get2d(ctx, f, host, m, n, x0, y0);
        const size_t es = (size_t)f.es;
        std::vector<char> col((size_t)n * es);
        HIPCHK(hipMemcpyAsync(col.data(), f.p0 + (ptrdiff_t)y0 * es, col.size(), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        char *o = (char *)host;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < m; ++i) std::memcpy(o + ((size_t)j * m + i) * es, &col[(size_t)j * es], es);
        return SWM_OK;
    };
    if ((rc = get(ctx->H, 1, H, c.nx + 2, c.ny + 2, 0, 0)) || (rc = get(ctx->Fx, ctx->Fx_sx, Fx, ctx->nux, ctx->nuy, 1 - ctx->ep, 1)) ||
        (rc = get(ctx->Fy, ctx->Fy_sx, Fy, ctx->nvx, ctx->nvy, 1, 1)) || (rc = get(ctx->fq, ctx->fq_sx, f_q, c.nx + 1, c.ny + 1, 0, 0)))
        return rc;
    if (eta_ref && c.continuity == SWM_CONT_RELAX && (rc = get(ctx->eta_ref, 1, eta_ref, c.nx, c.ny, 1, 1))) return rc;
    if (Feta && c.continuity == SWM_CONT_FORCING && (rc = get(ctx->Feta, 1, Feta, c.nx, c.ny, 1, 1))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return SWM_OK;
}

int swm_set_static_linear(swm_ctx *ctx, const void *f_u, const void *f_v) {
    if (!ctx) return SWM_E_ARG;
    if (ctx->c.dynamics != SWM_LINEAR) return ctx->fail(SWM_E_STATE, "f_u, f_v are read by the linear dynamics only");
    if (!f_u || !f_v) return ctx->fail(SWM_E_ARG, "f_u and f_v are required");
    if (ctx->have_fuv) return ctx->fail(SWM_E_STATE, "f_u, f_v already set");
    int rc;
    // f multiplies the velocities carried scaled, so it is not scaled itself
    if ((rc = put_static(ctx, ctx->f_u, &ctx->fu_sx, &ctx->fu_sy, f_u, ctx->nux, ctx->nuy, 1 - ctx->ep, 1))) return rc;
    if ((rc = put_static(ctx, ctx->f_v, &ctx->fv_sx, &ctx->fv_sy, f_v, ctx->nvx, ctx->nvy, 1, 1))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->have_fuv = true;
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

// host arrays of velocities carried scaled: ×1/s in their Tprog storage (exact)
static void unscale_host(const swm_ctx *ctx, void *a, size_t n) {
    if (!a || ctx->c.scale_log2 == 0) return;
    if (ctx->es == 8) for (size_t k = 0; k < n; ++k) ((double *)a)[k] *= ctx->sci;
    else for (size_t k = 0; k < n; ++k) ((float *)a)[k] *= (float)ctx->sci;
}

int swm_upload_prog(swm_ctx *ctx, const void *u, const void *v, const void *eta, const void *sst) {
    if (!ctx) return SWM_E_ARG;
    if (!u || !v || !eta || !sst) return ctx->fail(SWM_E_ARG, "null array");
    int rc = check_never_written(ctx, u, v);
    if (rc) return rc;
    const swm_config &c = ctx->c;
    const int pu = ctx->pu, pe = ctx->pe;
    if ((rc = put2d(ctx, ctx->u[pu], u, ctx->nux + 4, ctx->nuy + 4, -1 - ctx->ep, -1))) return rc;
    if ((rc = put2d(ctx, ctx->v[pu], v, ctx->nvx + 4, ctx->nvy + 4, -1, -1))) return rc;
    if ((rc = put2d(ctx, ctx->e[pe], eta, c.nx + 2, c.ny + 2, 0, 0))) return rc;
    if ((rc = put2d(ctx, ctx->sst[ctx->rsst], sst, c.nx + 2, c.ny, 0, 1))) return rc;
    rc = fmt_ops(ctx->fmt).upload_fix(ctx, pu, pe);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->have_prog = true;
    ctx->advcor_pending = (adv_form(c) & ADV_KEEP) != 0;
    return SWM_OK;
}

int swm_download_prog(swm_ctx *ctx, void *u, void *v, void *eta, void *sst) {
    if (!ctx) return SWM_E_ARG;
    if (!ctx->have_prog) return ctx->fail(SWM_E_STATE, "no prognostic state uploaded");
    const swm_config &c = ctx->c;
    int rc;
    const int pu = ctx->pu, pe = ctx->pe;
    if (u && (rc = get2d(ctx, ctx->u[pu], u, ctx->nux + 4, ctx->nuy + 4, -1 - ctx->ep, -1))) return rc;
    if (v && (rc = get2d(ctx, ctx->v[pu], v, ctx->nvx + 4, ctx->nvy + 4, -1, -1))) return rc;
    if (eta && (rc = get2d(ctx, ctx->e[pe], eta, c.nx + 2, c.ny + 2, 0, 0))) return rc;
    if (sst && (rc = get2d(ctx, ctx->sst[ctx->rsst], sst, c.nx + 2, c.ny, 0, 1))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    unscale_host(ctx, u, (size_t)(ctx->nux + 4) * (ctx->nuy + 4));
    unscale_host(ctx, v, (size_t)(ctx->nvx + 4) * (ctx->nvy + 4));
    return SWM_OK;
}

// the fields a step reads are all set
static int step_ready(swm_ctx *ctx) {
    if (!ctx->have_static || !ctx->have_prog) return ctx->fail(SWM_E_STATE, "set_static and upload_prog first");
    if (ctx->c.dynamics == SWM_LINEAR && !ctx->have_fuv)
        return ctx->fail(SWM_E_STATE, "the linear dynamics need f_u, f_v: swm_set_static_linear first");
    return SWM_OK;
}

int swm_step(swm_ctx *ctx, int64_t nsteps, int64_t i0, int64_t t0, int64_t *steps_done, int64_t *nan_step) {
    if (!ctx || !steps_done || !nan_step) return SWM_E_ARG;
    if (int rc = step_ready(ctx)) return rc;
    if (nsteps < 0 || i0 < 1) return ctx->fail(SWM_E_ARG, "nsteps >= 0 and i0 >= 1 required");
    return fmt_ops(ctx->fmt).run_steps(ctx, nsteps, i0, t0, steps_done, nan_step);
}

int swm_stages(swm_ctx *ctx, int32_t n_stages, int64_t i0, int64_t t0) {
    if (!ctx) return SWM_E_ARG;
    if (int rc = step_ready(ctx)) return rc;
    const swm_config &c = ctx->c;
    const int total = c.time_scheme == SWM_RK ? c.RKo : c.time_scheme == SWM_SSPRK2 ? c.RKs
                    : c.time_scheme == SWM_SSPRK3 ? c.ssp_s : 4;
    if (n_stages < 1 || n_stages > total) return ctx->fail(SWM_E_ARG, "n_stages %d out of 1..%d", n_stages, total);
    (void)i0;
    int rc = fmt_ops(ctx->fmt).stages(ctx, n_stages, t0);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return SWM_OK;
}

int swm_ghost(swm_ctx *ctx) {
    if (!ctx) return SWM_E_ARG;
    if (int rc = step_ready(ctx)) return rc;
    if (int rc = fmt_ops(ctx->fmt).ghost(ctx)) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return SWM_OK;
}

int swm_dissipation(swm_ctx *ctx) {
    if (!ctx) return SWM_E_ARG;
    if (int rc = step_ready(ctx)) return rc;
    if (int rc = fmt_ops(ctx->fmt).dissipation(ctx)) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return SWM_OK;
}

int swm_tracer(swm_ctx *ctx, int64_t i) {
    if (!ctx) return SWM_E_ARG;
    if (int rc = step_ready(ctx)) return rc;
    if (i < 1) return ctx->fail(SWM_E_ARG, "step i >= 1 required");
    if (int rc = fmt_ops(ctx->fmt).tracer(ctx, i)) return rc;
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

int swm_download_rk(swm_ctx *ctx, void *u0, void *v0, void *eta0, void *u1, void *v1, void *eta1) {
    if (!ctx) return SWM_E_ARG;
    if (ctx->rk0.uv < 0 || ctx->rk1.uv < 0) return ctx->fail(SWM_E_STATE, "call swm_stages first");
    const swm_config &c = ctx->c;
    int rc;
    const int mu = ctx->nux + 4, nu = ctx->nuy + 4, mv = ctx->nvx + 4, nv = ctx->nvy + 4;
    const Slots s0 = ctx->rk0, s1 = ctx->rk1;
    auto get = [&](const Field &f, void *h, int m, int n, int x0, int y0, bool t) {
        return t ? get2d_widen(ctx, f, h, m, n, x0, y0) : get2d(ctx, f, h, m, n, x0, y0);
    };
    if (u0 && (rc = get(ctx->fu(s0), u0, mu, nu, -1 - ctx->ep, -1, s0.t))) return rc;
    if (v0 && (rc = get(ctx->fv(s0), v0, mv, nv, -1, -1, s0.t))) return rc;
    if (eta0 && (rc = get(ctx->fe(s0), eta0, c.nx + 2, c.ny + 2, 0, 0, s0.t))) return rc;
    if (u1 && (rc = get(ctx->fu(s1), u1, mu, nu, -1 - ctx->ep, -1, s1.t))) return rc;
    if (v1 && (rc = get(ctx->fv(s1), v1, mv, nv, -1, -1, s1.t))) return rc;
    if (eta1 && (rc = get(ctx->fe(s1), eta1, c.nx + 2, c.ny + 2, 0, 0, s1.t))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (void *a : {u0, u1}) unscale_host(ctx, a, (size_t)mu * nu);
    for (void *a : {v0, v1}) unscale_host(ctx, a, (size_t)mv * nv);
    return SWM_OK;
}

int swm_time_kernel(swm_ctx *ctx, int32_t which, int32_t reps, double *ms_avg) {
    if (!ctx || !ms_avg) return SWM_E_ARG;
    if (int rc = step_ready(ctx)) return rc;
    if (reps < 1) return ctx->fail(SWM_E_ARG, "reps must be >= 1");
    return fmt_ops(ctx->fmt).time_kernel(ctx, which, reps, ms_avg);
}

int swm_time_steps(swm_ctx *ctx, int64_t nsteps, int64_t i0, int64_t t0, double *ms_avg, int32_t *counts) {
    if (!ctx || !ms_avg || !counts) return SWM_E_ARG;
    if (int rc = step_ready(ctx)) return rc;
    if (nsteps < 1 || i0 < 1) return ctx->fail(SWM_E_ARG, "nsteps >= 1 and i0 >= 1 required");
    StepTrace tr;
    ctx->strace = &tr;
    int64_t done = 0, nan = 0;
    int rc = fmt_ops(ctx->fmt).run_steps(ctx, nsteps, i0, t0, &done, &nan);
    ctx->strace = nullptr;
    const bool synced = hipStreamSynchronize(ctx->stream) == hipSuccess;
    for (int k = 0; k < KT_KINDS; ++k) {
        double sum = 0.0;
        int n = 0;
        for (size_t q = 0; q + 1 < tr.ev[k].size(); q += 2) {
            float ms = 0.0f;
            if (synced && hipEventElapsedTime(&ms, tr.ev[k][q], tr.ev[k][q + 1]) == hipSuccess) {
                sum += ms;
                ++n;
            }
        }
        for (hipEvent_t e : tr.ev[k]) (void)hipEventDestroy(e);
        ms_avg[k] = n ? sum / n : 0.0;
        counts[k] = n;
    }
    if (rc) return rc;
    if (!synced) return ctx->fail(SWM_E_HIP, "stream synchronize");
    if (nan) return ctx->fail(SWM_E_STATE, "NaN at step %lld", (long long)nan);
    return SWM_OK;
}

int swm_reduce(swm_ctx *ctx, double *mass, double *energy) {
    if (!ctx || !mass || !energy) return SWM_E_ARG;
    if (!ctx->have_static || !ctx->have_prog) return ctx->fail(SWM_E_STATE, "set_static and upload_prog first");
    const swm_config &c = ctx->c;
    const int pu = ctx->pu, pe = ctx->pe;
    int rc = fmt_ops(ctx->fmt).reduce_launch(ctx, pu, pe, ctx->dpart);
    if (rc) return rc;
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

int swm_reduce_noalloc(swm_ctx *ctx, double *mass, double *energy) {
    if (!ctx || !mass || !energy) return SWM_E_ARG;
    if (!ctx->have_static || !ctx->have_prog) return ctx->fail(SWM_E_STATE, "set_static and upload_prog first");
    const swm_config &c = ctx->c;
    if (c.ny > NOALLOC_MAX_ROWS)
        return ctx->fail(SWM_E_ARG, "ny = %d exceeds the %d rows the static scratch of swm_reduce_noalloc holds; use swm_reduce",
                         c.ny, NOALLOC_MAX_ROWS);
    double *s = (ctx->dev >= 0 && ctx->dev < MAX_DEVICES) ? g_scratch[ctx->dev] : nullptr;
    if (!s) return ctx->fail(SWM_E_STATE, "no scratch address for device %d", ctx->dev);
    int rc = fmt_ops(ctx->fmt).reduce_launch(ctx, ctx->pu, ctx->pe, s);
    if (rc) return rc;
    double *sums = s + 2 * NOALLOC_MAX_ROWS;
    hipLaunchKernelGGL(reduce_rows_kernel, dim3(1), dim3(64), 0, ctx->stream, (const double *)s, c.ny, sums);
    HIPCHK(hipGetLastError());
    double out[2] = {0.0, 0.0};
    HIPCHK(hipMemcpyAsync(out, sums, sizeof out, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    *mass = out[0];
    *energy = out[1];
    return SWM_OK;
}

int swm_ens_create(int32_t n, swm_ctx *const *members, swm_ens **out) {
    if (!out || n < 1 || !members) return SWM_E_ARG;
    *out = nullptr;
    swm_ctx *m0 = members[0];
    if (!m0) return SWM_E_ARG;
    const swm_config &c0 = m0->c;
    for (int k = 0; k < n; ++k) {
        swm_ctx *ctx = members[k];
        if (!ctx) return SWM_E_ARG;
        const swm_config &c = ctx->c;
        const char *why = nullptr;
        if (!ctx->have_static || !ctx->have_prog) why = "static fields and a state must be uploaded";
        else if (c.dynamics == SWM_LINEAR && !ctx->have_fuv) why = "the linear dynamics need f_u, f_v (swm_set_static_linear)";
        else if (ctx->dev != m0->dev) why = "all members must live on one device";
        else if (c.T != c0.T || c.Tprog != c0.Tprog || c.Tcomm != c0.Tcomm) why = "number formats differ";
        else if (c.bc != c0.bc || c.time_scheme != SWM_RK || c.RKo != c0.RKo) why = "boundary condition or time scheme differs (classic RK only)";
        else if (c.bottom_drag != c0.bottom_drag || c.diffusion != c0.diffusion || c.continuity != c0.continuity) why = "drag, diffusion or continuity differs";
        else if (adv_form(c) != adv_form(c0)) why = "right-hand side form differs (scheme, dynamics, nstep_advcor)";
        else if (c.nstep_advcor != c0.nstep_advcor) why = "nstep_advcor differs";
        else if (c.nstep_diff != 1) why = "nstep_diff must be 1";
        else if (c.compensated || c.scale_log2 != c0.scale_log2) why = "compensated members, or members of different scale";
        else if (ctx->fuse != m0->fuse) why = "stage fusion setting differs";
        if (why) {
            fprintf(stderr, "swm_ens_create: member %d: %s\n", k, why);
            return SWM_E_UNSUPPORTED;
        }
        for (int j = 0; j < k; ++j)
            if (members[j] == ctx) return SWM_E_ARG;
    }
    swm_ens *E = new swm_ens();
    E->m.assign(members, members + n);
    E->fmt = m0->fmt;
    E->per = c0.bc == SWM_PERIODIC;
    if (hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&E->stream2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&E->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&E->join, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        if (E->stream) (void)hipStreamDestroy(E->stream);
        if (E->stream2) (void)hipStreamDestroy(E->stream2);
        if (E->fork) (void)hipEventDestroy(E->fork);
        if (E->join) (void)hipEventDestroy(E->join);
        delete E;
        return SWM_E_HIP;
    }
    for (swm_ctx *c : E->m) c->ens_refs += 1;
    // periodic members run the fused pairs that wrap around the seam only if every member's
    // seam allows it, else all run the form with the seam strips on their own kernel (one
    // kernel sequence for all members); the two forms give the same results
    bool seam_all = true;
    for (swm_ctx *c : E->m) seam_all = seam_all && c->seam_ok;
    if (!seam_all)
        for (swm_ctx *c : E->m) c->seam_ok = false;
    *out = E;
    return SWM_OK;
}

int swm_ens_destroy(swm_ens *E) {
    if (!E) return SWM_OK;
    if (E->stream) (void)hipStreamSynchronize(E->stream);
    if (E->stream2) (void)hipStreamSynchronize(E->stream2);
    if (E->exec) (void)hipGraphExecDestroy(E->exec);
    if (E->graph) (void)hipGraphDestroy(E->graph);
    for (auto &b : E->buf) if (b.dev) (void)hipFree(b.dev);
    if (E->pdev) (void)hipFree(E->pdev);
    if (E->stream) (void)hipStreamDestroy(E->stream);
    if (E->stream2) (void)hipStreamDestroy(E->stream2);
    if (E->fork) (void)hipEventDestroy(E->fork);
    if (E->join) (void)hipEventDestroy(E->join);
    for (swm_ctx *c : E->m) c->ens_refs -= 1;
    delete E;
    return SWM_OK;
}

const char *swm_ens_last_error(const swm_ens *E) { return E ? E->err.c_str() : "null ensemble"; }

int swm_ens_tune(swm_ens *E, int32_t reps, int32_t *rows_stage, int32_t *rows_fused, int32_t *rows_diss) {
    if (!E || reps < 1) return SWM_E_ARG;
    for (swm_ctx *c : E->m) (void)hipStreamSynchronize(c->stream);
    StreamSwap sw(E);
    int Ls = 0, Lf = 0, Ld = 0;
    int rc = fmt_ops(E->fmt).ens_tune(E, reps, &Ls, &Lf, &Ld);
    if (rc) return rc;
    if (rows_stage) *rows_stage = Ls;
    if (rows_fused) *rows_fused = Lf;
    if (rows_diss) *rows_diss = Ld;
    return SWM_OK;
}

int swm_ens_step(swm_ens *E, int64_t nsteps, int64_t i0, int64_t *steps_done, int64_t *nan_steps, int64_t *graph_steps) {
    if (!E || !steps_done || !nan_steps || !graph_steps || nsteps < 0 || i0 < 1) return SWM_E_ARG;
    for (swm_ctx *c : E->m) (void)hipStreamSynchronize(c->stream);
    StreamSwap sw(E);
    return fmt_ops(E->fmt).ens_step(E, nsteps, i0, steps_done, nan_steps, graph_steps);
}

int swm_ens_set_persistent(swm_ens *E, int32_t on) {
    if (!E || on < 0 || on > 2) return SWM_E_ARG;
    E->persistent = on;
    return SWM_OK;
}

int swm_ens_phase_ms(const swm_ens *E, double *ms) {
    if (!E || !ms) return SWM_E_ARG;
    for (int k = 0; k < 4; ++k) ms[k] = E->phase_ms[k];
    return SWM_OK;
}

int swm_ens_persistent_steps(const swm_ens *E, int64_t *steps, const char **why_not) {
    if (!E || !steps) return SWM_E_ARG;
    *steps = E->persist_steps;
    if (why_not) *why_not = E->persist_why;
    return SWM_OK;
}

// ---- reverse mode

int swm_adj_forward(swm_ctx *ctx, int64_t nsteps, int64_t i0, int64_t t0, int32_t snapshots) {
    if (!ctx) return SWM_E_ARG;
    if (nsteps < 1 || i0 < 1) return ctx->fail(SWM_E_ARG, "nsteps >= 1 and i0 >= 1 required");
    if (int rc = adj_ok(ctx)) return rc;
    return fmt_ops(ctx->fmt).adj_forward(ctx, nsteps, i0, t0, snapshots);
}

int swm_adj_pullback(swm_ctx *ctx, const void *ybar_u, const void *ybar_v, const void *ybar_eta, const void *ybar_sst,
                     void *xbar_u, void *xbar_v, void *xbar_eta, void *xbar_sst, int64_t *forward_steps, double *sweep_ms) {
    if (!ctx) return SWM_E_ARG;
    if (!ybar_u || !ybar_v || !ybar_eta || !ybar_sst) return ctx->fail(SWM_E_ARG, "null cotangent array");
    if (!ctx->adj || !ctx->adj->recorded) return ctx->fail(SWM_E_STATE, "call swm_adj_forward first");
    int rc = adj_ok(ctx);
    if (!rc) rc = fmt_ops(ctx->fmt).adj_pullback(ctx, ybar_u, ybar_v, ybar_eta, ybar_sst, xbar_u, xbar_v, xbar_eta, xbar_sst);
    if (forward_steps) *forward_steps = ctx->adj->fwd_steps;
    if (sweep_ms) *sweep_ms = ctx->adj->sweep_ms;
    return rc;
}

int swm_adj_track_Fx(swm_ctx *ctx, int32_t on) {
    if (!ctx) return SWM_E_ARG;
    if (int rc = adj_ok(ctx)) return rc;
    ctx->adj_fx = on != 0;
    return SWM_OK;
}

int swm_adj_download_Fx(swm_ctx *ctx, void *xbar_Fx) {
    if (!ctx || !xbar_Fx) return SWM_E_ARG;
    if (int rc = adj_ok(ctx)) return rc;
    return fmt_ops(ctx->fmt).adj_download_fx(ctx, xbar_Fx);
}

int swm_adj_time(swm_ctx *ctx, int32_t which, int32_t reps, double *ms_avg) {
    if (!ctx || !ms_avg) return SWM_E_ARG;
    if (reps < 1) return ctx->fail(SWM_E_ARG, "reps must be >= 1");
    if (int rc = adj_ok(ctx)) return rc;
    return fmt_ops(ctx->fmt).adj_time(ctx, which, reps, ms_avg);
}

}  // extern "C"
