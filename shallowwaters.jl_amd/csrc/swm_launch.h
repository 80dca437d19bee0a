// swm_launch.h -- kernel launches of the step (Launch), the reverse mode (Adj) and batched
// ensembles (Ens) for one number-format bundle F.  Each swm_<format>.hip instantiates
// them for its bundle and exports the table of entry points the C ABI dispatches to.
#pragma once

#include "swm_host.h"

#include "k_adjoint.h"
#include "k_dissipation.h"
#include "k_misc.h"
#include "k_persist.h"
#include "k_setup.h"
#include "k_stage.h"

namespace swmh {

template <class F> struct Launch {
    using R = typename F::R;
    using Q = typename F::Q;
    using ST = typename F::ST;

    static int nstrips(const swm_ctx *ctx) { return (ctx->c.nx + 1 + WOUT - 1) / WOUT; }
    static int nstrips2(const swm_ctx *ctx) { return (ctx->c.nx + 1 + WOUT2 - 1) / WOUT2; }
    static int nseg(const swm_ctx *ctx, int L) { return (ctx->c.ny + L - 1) / L; }

    // Rows per wave segment of `kernel` (strips of `nstr`), for `waves_per_slot` resident
    // waves per wave slot the kernel's occupancy gives (not below lmin rows).  A segment
    // re-reads the halo rows it shares with its neighbours; the stage kernels read their top
    // halo at the start, while the segment above reads the same rows, so the re-read stays
    // in L2 and more, shorter segments pay.  Measured at 4096^2 on MI355X: about six waves
    // per slot for the stage kernels, five for the fused pairs, two for dissipation.
    static int rows_per_wave(const swm_ctx *ctx, const void *kernel, int nstr, int waves_per_slot, int lmin) {
        int cus = 0, bpc = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->dev) != hipSuccess || cus <= 0) cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, kernel, WAVE * WAVES_PER_BLOCK, 0) != hipSuccess || bpc <= 0)
            bpc = 1;
        const long slots = (long)cus * bpc * WAVES_PER_BLOCK * waves_per_slot;
        const long ny = ctx->c.ny;
        const long L = (ny * nstr + slots - 1) / slots;
        return (int)std::max<long>(lmin, std::min<long>(L, std::max<long>(ny, lmin)));
    }
    template <bool PER> static void segment_lengths(swm_ctx *ctx) {
        ctx->L = rows_per_wave(ctx, reinterpret_cast<const void *>(&stage_kernel<F, PER, M_RK_MID>), nstrips(ctx), 6, 4);
        ctx->Ld = rows_per_wave(ctx, reinterpret_cast<const void *>(&dissipation_kernel<F, PER, 0, true>), nstrips(ctx), 2, 8);
        ctx->Lf = rows_per_wave(ctx, reinterpret_cast<const void *>(&fused_kernel<F, M_RK_MID, M_RK_LAST, false>),
                                nstrips2(ctx), 5, 8);
        if (const char *e = getenv("SWM_ROWS_PER_WAVE_FUSED")) {     // rows per wave of the fused pairs, fixed
            const int l = atoi(e);
            if (l >= 8 && l <= 4096) ctx->Lf = l;
        }
    }

    static StaticRefs statics(const swm_ctx *ctx) {
        StaticRefs s;
        s.H = ctx->H.ref();
        s.Fx = ctx->Fx.ref(); s.Fx_2d = ctx->Fx_sx;
        s.Fy = ctx->Fy.ref(); s.Fy_2d = ctx->Fy_sx;
        s.fq = ctx->fq.ref(); s.fq_2d = ctx->fq_sx;
        s.eta_ref = ctx->eta_ref.ref();
        s.Feta = ctx->Feta.ref();
        s.fu = ctx->f_u.ref(); s.fu_2d = ctx->fu_sx;
        s.fv = ctx->f_v.ref(); s.fv_2d = ctx->fv_sx;
        return s;
    }
    static uint32_t ou(const swm_ctx *ctx, Slots s) { return s.uv < 0 ? 0u : ctx->fu(s).ref(); }
    static uint32_t ov(const swm_ctx *ctx, Slots s) { return s.uv < 0 ? 0u : ctx->fv(s).ref(); }
    static uint32_t oe(const swm_ctx *ctx, Slots s) { return s.e < 0 ? 0u : ctx->fe(s).ref(); }

    struct Times { R Fxt, Fyt, Fet; };
    static Times forcing_times(const swm_ctx *ctx, int64_t t) {     // Ftime, forcing.jl:215-217: T(sin(ω t))
        const swm_config &c = ctx->c;
        Times T;
        T.Fxt = mkR<R>(c.seasonal_wind_x ? round_host(std::sin(c.omegaFx * (double)t), c.T) : 1.0);
        T.Fyt = mkR<R>(c.seasonal_wind_y ? round_host(std::sin(c.omegaFy * (double)t), c.T) : 1.0);
        T.Fet = mkR<R>(c.continuity == SWM_CONT_FORCING ? round_host(std::sin(c.omegaFeta * (double)t), c.T) : 0.0);
        return T;
    }

    static StageArgs<F> args(swm_ctx *ctx, Slots in, Slots pr, Slots ac, Slots o1, Slots o2, const Coef<Q> &k,
                             const Times &T) {
        const swm_config &c = ctx->c;
        // every byte set (padding included): an ensemble compares the recorded arguments of
        // consecutive steps byte for byte to replay them as one graph
        StageArgs<F> A;
        std::memset((void *)&A, 0, sizeof A);
        A.sfirst = 0; A.sjump_at = 1 << 30; A.sjump = 0;
        A.ar = ctx->aref();
        A.in_u = ou(ctx, in); A.in_v = ov(ctx, in); A.in_e = oe(ctx, in);
        A.pr_u = ou(ctx, pr); A.pr_v = ov(ctx, pr); A.pr_e = oe(ctx, pr);
        A.ac_u = ou(ctx, ac); A.ac_v = ov(ctx, ac); A.ac_e = oe(ctx, ac);
        A.o1_u = ou(ctx, o1); A.o1_v = ov(ctx, o1); A.o1_e = oe(ctx, o1);
        A.o2_u = ou(ctx, o2); A.o2_v = ov(ctx, o2); A.o2_e = oe(ctx, o2);
        A.qend = ctx->qend.ref();
        A.qs = ctx->qs.ref(); A.ks = ctx->ks.ref();
        A.st = statics(ctx);
        A.k = k;
        A.Fxt = T.Fxt; A.Fyt = T.Fyt; A.Fet = T.Fet;
        // scaled velocities: PV weights /s, Bernoulli ½ /s and g·s (powers of two, exact)
        A.g = mkR<R>(c.g * ctx->sc); A.gamma = mkR<R>(c.gamma); A.k12 = mkR<R>(round_host(c.adv_scheme == SWM_ADV_SADOURNY ? 0.5 : 1.0 / 12.0, c.T) * ctx->sci);
        A.halfp = mkR<R>(0.5 * ctx->sci);
        A.oma = (Q)c.one_minus_alpha;
        A.sinv = (Q)ctx->sci;
        A.scaled = c.scale_log2 != 0;
        A.c_u = ctx->cu.ref(); A.c_v = ctx->cv.ref(); A.c_e = ctx->ce.ref();
        A.cont = c.continuity; A.cm = c.Tcomm;
        A.nx = c.nx; A.ny = c.ny; A.P = ctx->P; A.L = ctx->L;
        A.nstrips = nstrips(ctx); A.nseg = nseg(ctx, ctx->L);
        return A;
    }

    // Every buffer a mode reads or writes must name a slot (no launch through a null base),
    // and the stage input / first output live in the pool the mode's formats expect.
    static bool slots_ok(const swm_ctx *ctx, int m, Slots in, Slots pr, Slots ac, Slots o1, Slots o2, bool cont) {
        const bool rk = m <= M_RK_LAST || m == M_RKC_FIRST || m == M_RKC_LAST;
        const bool need_pr = m == M_RK_MID || m == M_SSP2_LAST || m == M_S4_3 || m == M_RKC_LAST;
        const bool need_ac = m == M_RK_MID || m == M_RK_LAST || m == M_SSP3_KN || m == M_RKC_LAST;
        const bool out1 = m != M_RK_LAST && m != M_RKC_LAST;
        const bool out2 = m == M_RK_FIRST || m == M_RK_MID || m == M_RK_LAST || m == M_SSP2_LAST || (m == M_S4_3 && !cont) ||
                          m == M_RKC_FIRST || m == M_RKC_LAST;
        if (m == M_RKC_LAST && !ctx->cu.base) return false;
        auto ok = [ctx](Slots s, bool uv, bool e, bool tpool) {
            const int n = tpool ? 2 : ctx->nslots;
            if (s.t != tpool) return false;
            return (!uv || (s.uv >= 0 && s.uv < n)) && (!e || (s.e >= 0 && s.e < n));
        };
        const bool uv = !cont, e = rk || cont;            // which of (u,v) and η the kernel touches
        const bool tin = ctx->mixed && m != M_RK_FIRST && m != M_RKC_FIRST;   // stage input u1 in the T pool
        const bool tout = ctx->mixed;                     // first output u1 in the T pool
        if (!ok(in, uv, e, tin)) return false;
        if (need_pr && !ok(pr, uv, e, false)) return false;
        if (need_ac && !ok(ac, uv, e, false)) return false;
        if (out1 && !ok(o1, uv, e, tout)) return false;
        if (out2 && !ok(o2, uv, e, false)) return false;
        return true;
    }

    // one stage kernel (momentum, and continuity for the RK modes) in mode m
    static int stage(swm_ctx *ctx, int m, Slots in, Slots pr, Slots ac, Slots o1, Slots o2, const Coef<Q> &k,
                     const Times &T) {
        const swm_config &c = ctx->c;
        const bool per = c.bc == SWM_PERIODIC;
        if (!slots_ok(ctx, m, in, pr, ac, o1, o2, false)) return ctx->fail(SWM_E_STATE, "stage mode %d: buffer slot missing", m);
        StageArgs<F> A = args(ctx, in, pr, ac, o1, o2, k, T);
        if (ctx->rec) {
            ctx->rec->add(0, m, -1, &A, sizeof A);
            return SWM_OK;
        }
        const int adv = adv_form(c);
        if (per && adv == ADV_AH) {
            const dim3 qg((c.ny + 1 + 255) / 256), qb(256);
            if (m == M_RK_FIRST || m == M_RKC_FIRST) hipLaunchKernelGGL((qend_kernel<F, Q>), qg, qb, 0, ctx->stream, A);
            else hipLaunchKernelGGL((qend_kernel<F, ST>), qg, qb, 0, ctx->stream, A);
        }
        const dim3 grid(nblocks_waves(A.nstrips * A.nseg)), block(WAVE * WAVES_PER_BLOCK);
        const bool first = m == M_RK_FIRST || m == M_RKC_FIRST, last = m == M_RK_LAST || m == M_RKC_LAST;
        const KTrace kt(ctx, first ? KT_STAGE_FIRST : last ? KT_STAGE_LAST : KT_STAGE_MID);
#define SWM_KA(M_, ADV_)                                                                                   \
    if (per) hipLaunchKernelGGL((stage_kernel<F, true, M_, ADV_>), grid, block, 0, ctx->stream, A);        \
    else hipLaunchKernelGGL((stage_kernel<F, false, M_, ADV_>), grid, block, 0, ctx->stream, A);
#define SWM_K(M_)                                                                                          \
    case M_:                                                                                               \
        if (adv == ADV_SAD) { SWM_KA(M_, ADV_SAD) }                                                        \
        else if (adv == ADV_LIN) { SWM_KA(M_, ADV_LIN) }                                                   \
        else if (adv == (ADV_AH | ADV_KEEP)) { SWM_KA(M_, ADV_AH | ADV_KEEP) }                             \
        else if (adv == (ADV_SAD | ADV_KEEP)) { SWM_KA(M_, ADV_SAD | ADV_KEEP) }                           \
        else { SWM_KA(M_, ADV_AH) }                                                                        \
        break;
        if constexpr (F::mixed) {
            switch (m) {
                SWM_K(M_RK_FIRST) SWM_K(M_RK_MID) SWM_K(M_RK_LAST) SWM_K(M_RKC_FIRST) SWM_K(M_RKC_LAST)
            default: return ctx->fail(SWM_E_UNSUPPORTED, "stage mode %d with mixed formats", m);
            }
        } else {
            switch (m) {
                SWM_K(M_RK_FIRST) SWM_K(M_RK_MID) SWM_K(M_RK_LAST) SWM_K(M_RKC_FIRST) SWM_K(M_RKC_LAST) SWM_K(M_SSP2)
                SWM_K(M_SSP2_LAST) SWM_K(M_SSP3)
                SWM_K(M_SSP3_KN) SWM_K(M_S4) SWM_K(M_S4_3)
            default: return ctx->fail(SWM_E_ARG, "stage mode %d", m);
            }
        }
#undef SWM_K
#undef SWM_KA
        HIPCHK(hipGetLastError());
        return SWM_OK;
    }

    // q and KEu+KEv kept for nstep_advcor > 0, formed from the state in slots s (Tprog
    // storage): wq also q; init with the initial state's thickness
    static int advcor(swm_ctx *ctx, Slots s, bool wq, bool init) {
        const swm_config &c = ctx->c;
        const Times T{mkR<R>(1.0), mkR<R>(1.0), mkR<R>(0.0)};
        const StageArgs<F> A = args(ctx, s, Slots{}, Slots{}, Slots{}, Slots{}, Coef<Q>{}, T);
        if (ctx->rec) {
            ctx->rec->add(6, (init ? 2 : 0) | (wq ? 1 : 0), -1, &A, sizeof A);
            return SWM_OK;
        }
        const dim3 g((c.nx + 2 + 255) / 256, c.ny + 2), b(256);
        if (init) hipLaunchKernelGGL((advcor_kernel<F, Q, true, true>), g, b, 0, ctx->stream, A);
        else if (wq) hipLaunchKernelGGL((advcor_kernel<F, Q, true, false>), g, b, 0, ctx->stream, A);
        else hipLaunchKernelGGL((advcor_kernel<F, Q, false, false>), g, b, 0, ctx->stream, A);
        HIPCHK(hipGetLastError());
        return SWM_OK;
    }

    // two consecutive classic RK stages in one kernel (nonperiodic): mode ma then mb
    static bool fusable(const swm_ctx *ctx, int ma, int mb) { return ctx->fuse && can_fuse(ctx, ma, mb); }
    static bool can_fuse(const swm_ctx *ctx, int ma, int mb) {
        const swm_config &c = ctx->c;
        if (c.time_scheme != SWM_RK) return false;
        // periodic: a seam the lanes can wrap around (one launch), or the seam strips on their
        // own kernel; an ensemble's batched pairs take the first kind only
        if (adv_form(c) == ADV_AH) return true;
        // Sadourny, the linear dynamics, kept q and KE: the pairs of RK4 with T = Tprog, no
        // continuity static, on a seam the lanes can wrap around
        return !ctx->mixed && c.continuity == SWM_CONT_ITSELF && (c.bc == SWM_NONPERIODIC || ctx->seam_ok) &&
               ((ma == M_RK_FIRST && mb == M_RK_MID) || (ma == M_RK_MID && mb == M_RK_LAST));
    }
    static int fused(swm_ctx *ctx, int ma, int mb, Slots in, Slots pr, Slots ac, Slots o1, Slots o2, const Coef<Q> &ka,
                     const Coef<Q> &kb, const Times &T) {
        // the first stage's outputs stay in registers: check what the pair reads and writes
        auto ok = [](Slots s, bool tpool) {
            const int n = tpool ? 2 : 4;
            return s.t == tpool && s.uv >= 0 && s.uv < n && s.e >= 0 && s.e < n;
        };
        const bool afirst = ma == M_RK_FIRST || ma == M_RKC_FIRST, blast = mb == M_RK_LAST || mb == M_RKC_LAST;
        const bool good = ok(in, ctx->mixed && !afirst) && ok(pr, false) && (afirst || ok(ac, false)) &&
                          (blast || ok(o1, ctx->mixed)) && ok(o2, false) && (mb != M_RKC_LAST || ctx->cu.base);
        if (!good) return ctx->fail(SWM_E_STATE, "fused stages %d+%d: buffer slot missing", ma, mb);
        const bool per = ctx->c.bc == SWM_PERIODIC;
        if (!can_fuse(ctx, ma, mb)) return ctx->fail(SWM_E_UNSUPPORTED, "stages %d+%d do not fuse in this configuration", ma, mb);
        const int adv = adv_form(ctx->c);
        StageArgs<F> A = args(ctx, in, pr, ac, o1, o2, ka, T);
        A.k2 = kb;
        A.nstrips = nstrips2(ctx);
        A.L = ctx->Lf;
        A.nseg = nseg(ctx, ctx->Lf);
        if (per && !ctx->seam_ok) return fused_seam(ctx, ma, mb, A);
        if (ctx->rec) {
            ctx->rec->add(1, ma, mb, &A, sizeof A);
            return SWM_OK;
        }
        const dim3 grid(nblocks_waves(A.nstrips * A.nseg)), block(WAVE * WAVES_PER_BLOCK);
        const bool cs = ctx->c.continuity != SWM_CONT_ITSELF;
        const KTrace kt(ctx, ma == M_RK_FIRST || ma == M_RKC_FIRST ? KT_FUSED12 : KT_FUSED34);
        if (adv != ADV_AH) {
            if constexpr (!F::mixed) {
#define SWM_FV(MA_, MB_, ADV_)                                                                             \
    if (per) hipLaunchKernelGGL((fused_kernel<F, MA_, MB_, false, true, ADV_>), grid, block, 0, ctx->stream, A); \
    else hipLaunchKernelGGL((fused_kernel<F, MA_, MB_, false, false, ADV_>), grid, block, 0, ctx->stream, A)
#define SWM_FVA(ADV_)                                                                                      \
    if (ma == M_RK_FIRST) SWM_FV(M_RK_FIRST, M_RK_MID, ADV_);                                              \
    else SWM_FV(M_RK_MID, M_RK_LAST, ADV_)
                if (adv == ADV_SAD) { SWM_FVA(ADV_SAD); }
                else if (adv == ADV_LIN) { SWM_FVA(ADV_LIN); }
                else if (adv == (ADV_AH | ADV_KEEP)) { SWM_FVA(ADV_AH | ADV_KEEP); }
                else if (adv == (ADV_SAD | ADV_KEEP)) { SWM_FVA(ADV_SAD | ADV_KEEP); }
                else return ctx->fail(SWM_E_ARG, "right-hand side form %d", adv);
#undef SWM_FVA
#undef SWM_FV
                HIPCHK(hipGetLastError());
                return SWM_OK;
            }
            return ctx->fail(SWM_E_UNSUPPORTED, "fused stages of this right-hand side with mixed formats");
        }
#define SWM_FK(MA_, MB_)                                                                                   \
    if (per) {                                                                                             \
        if (cs) hipLaunchKernelGGL((fused_kernel<F, MA_, MB_, true, true>), grid, block, 0, ctx->stream, A); \
        else hipLaunchKernelGGL((fused_kernel<F, MA_, MB_, false, true>), grid, block, 0, ctx->stream, A);   \
    } else if (cs) hipLaunchKernelGGL((fused_kernel<F, MA_, MB_, true>), grid, block, 0, ctx->stream, A);  \
    else hipLaunchKernelGGL((fused_kernel<F, MA_, MB_, false>), grid, block, 0, ctx->stream, A)
        if (ma == M_RK_FIRST && mb == M_RK_MID) SWM_FK(M_RK_FIRST, M_RK_MID);
        else if (ma == M_RK_MID && mb == M_RK_LAST) SWM_FK(M_RK_MID, M_RK_LAST);
        else if (ma == M_RK_FIRST && mb == M_RK_LAST) SWM_FK(M_RK_FIRST, M_RK_LAST);
        else if (ma == M_RKC_FIRST && mb == M_RK_MID) SWM_FK(M_RKC_FIRST, M_RK_MID);
        else if (ma == M_RK_MID && mb == M_RKC_LAST) SWM_FK(M_RK_MID, M_RKC_LAST);
        else if (ma == M_RKC_FIRST && mb == M_RKC_LAST) SWM_FK(M_RKC_FIRST, M_RKC_LAST);
        else return ctx->fail(SWM_E_ARG, "fused stages %d+%d", ma, mb);
#undef SWM_FK
        HIPCHK(hipGetLastError());
        return SWM_OK;
    }

    // A fused pair of the adjoint's taped step: as fused(), and the first stage's output is
    // also stored, in slots tp (T = Tprog, Arakawa-Hsu, a seam the lanes wrap around).
    static bool can_tape(const swm_ctx *ctx, int ma, int mb) {
        return !F::mixed && ctx->fuse && can_fuse(ctx, ma, mb) && adv_form(ctx->c) == ADV_AH &&
               (ctx->c.bc == SWM_NONPERIODIC || ctx->seam_ok) && !ctx->rec;
    }
    static int fused_tape(swm_ctx *ctx, int ma, int mb, Slots in, Slots pr, Slots ac, Slots o1, Slots o2, Slots tp,
                          const Coef<Q> &ka, const Coef<Q> &kb, const Times &T) {
        auto ok = [&](Slots s) { return !s.t && s.uv >= 0 && s.uv < ctx->nslots && s.e >= 0 && s.e < ctx->nslots; };
        const bool afirst = ma == M_RK_FIRST, blast = mb == M_RK_LAST;
        if (!can_tape(ctx, ma, mb) || !ok(in) || !ok(pr) || !(afirst || ok(ac)) || !(blast || ok(o1)) || !ok(o2) || !ok(tp))
            return ctx->fail(SWM_E_STATE, "taped fused stages %d+%d: not available here", ma, mb);
        StageArgs<F> A = args(ctx, in, pr, ac, o1, o2, ka, T);
        A.k2 = kb;
        A.nstrips = nstrips2(ctx);
        A.L = ctx->Lf;
        A.nseg = nseg(ctx, ctx->Lf);
        A.tp_u = ou(ctx, tp); A.tp_v = ov(ctx, tp); A.tp_e = oe(ctx, tp);
        const dim3 grid(nblocks_waves(A.nstrips * A.nseg)), block(WAVE * WAVES_PER_BLOCK);
        const bool per = ctx->c.bc == SWM_PERIODIC, cs = ctx->c.continuity != SWM_CONT_ITSELF;
        if constexpr (!F::mixed) {
#define SWM_FT(MA_, MB_)                                                                                                \
    if (per) {                                                                                                          \
        if (cs) hipLaunchKernelGGL((fused_kernel<F, MA_, MB_, true, true, ADV_TAPE>), grid, block, 0, ctx->stream, A);   \
        else hipLaunchKernelGGL((fused_kernel<F, MA_, MB_, false, true, ADV_TAPE>), grid, block, 0, ctx->stream, A);     \
    } else if (cs) hipLaunchKernelGGL((fused_kernel<F, MA_, MB_, true, false, ADV_TAPE>), grid, block, 0, ctx->stream, A); \
    else hipLaunchKernelGGL((fused_kernel<F, MA_, MB_, false, false, ADV_TAPE>), grid, block, 0, ctx->stream, A)
            if (ma == M_RK_FIRST && mb == M_RK_MID) SWM_FT(M_RK_FIRST, M_RK_MID);
            else if (ma == M_RK_MID && mb == M_RK_LAST) SWM_FT(M_RK_MID, M_RK_LAST);
            else if (ma == M_RK_FIRST && mb == M_RK_LAST) SWM_FT(M_RK_FIRST, M_RK_LAST);
            else return ctx->fail(SWM_E_ARG, "taped fused stages %d+%d", ma, mb);
#undef SWM_FT
        }
        HIPCHK(hipGetLastError());
        return SWM_OK;
    }

    // A periodic fused pair whose seam the lanes cannot wrap around: the strips clear of
    // the seam on the wrapping kernel (no lane of theirs wraps), and beside them, on the
    // context's second stream, the strips at the seam on fused_seam_kernel.  The two write
    // disjoint columns and read only what the launches before the pair wrote.
    static int fused_seam(swm_ctx *ctx, int ma, int mb, StageArgs<F> A) {
        const swm_config &c = ctx->c;
        const int ns = A.nstrips, wo = WAVE - 2 * HALO2;
        // a strip is clear of the seam if all its lanes lie in the columns 1..nx-1
        auto clear = [&](int s) { return wo * s - HALO2 >= 1 && wo * s + WAVE - HALO2 - 1 <= c.nx - 1; };
        int a = 0;
        while (a < ns && !clear(a)) ++a;
        int b = a;
        while (b < ns && clear(b)) ++b;
        const dim3 block(WAVE * WAVES_PER_BLOCK);
        const bool cs = c.continuity != SWM_CONT_ITSELF;
        StageArgs<F> I = A, E = A;
        I.nstrips = b - a; I.sfirst = a;
        E.nstrips = a + (ns - b); E.sfirst = 0; E.sjump_at = a; E.sjump = b - a;
        // Segments of 4 rows for the seam strips: their waves run alone on their SIMDs, so the
        // launch's time is the march of one segment (4 rows: 2048x1024 over the ridge 0.113 ms
        // per step against 0.126 with the pairs' own 8 rows, 0.215 with 32; 4096² 0.565
        // against 0.577).
        E.L = 4;
        E.nseg = nseg(ctx, E.L);
        if (ctx->rec) {                                      // an ensemble's plan: the two launches as items
            ctx->rec->add(4, ma, mb, &I, sizeof I);
            ctx->rec->add(5, ma, mb, &E, sizeof E);
            return SWM_OK;
        }
        hipStream_t s2 = I.nstrips > 0 ? ctx->stream2 : ctx->stream;
        if (s2 != ctx->stream) {                             // fork
            HIPCHK(hipEventRecord(ctx->fork, ctx->stream));
            HIPCHK(hipStreamWaitEvent(s2, ctx->fork, 0));
        }
#define SWM_FS(MA_, MB_)                                                                                       \
    {                                                                                                          \
        const dim3 g(nblocks_waves(E.nstrips * E.nseg));                                                       \
        if (cs) hipLaunchKernelGGL((fused_seam_kernel<F, MA_, MB_, true>), g, block, 0, s2, E);                \
        else hipLaunchKernelGGL((fused_seam_kernel<F, MA_, MB_, false>), g, block, 0, s2, E);                  \
    }                                                                                                          \
    if (I.nstrips > 0) {                                                                                       \
        const dim3 g(nblocks_waves(I.nstrips * I.nseg));                                                       \
        if (cs) hipLaunchKernelGGL((fused_kernel<F, MA_, MB_, true, true>), g, block, 0, ctx->stream, I);      \
        else hipLaunchKernelGGL((fused_kernel<F, MA_, MB_, false, true>), g, block, 0, ctx->stream, I);        \
    }
        if (ma == M_RK_FIRST && mb == M_RK_MID) { SWM_FS(M_RK_FIRST, M_RK_MID) }
        else if (ma == M_RK_MID && mb == M_RK_LAST) { SWM_FS(M_RK_MID, M_RK_LAST) }
        else if (ma == M_RK_FIRST && mb == M_RK_LAST) { SWM_FS(M_RK_FIRST, M_RK_LAST) }
        else if (ma == M_RKC_FIRST && mb == M_RK_MID) { SWM_FS(M_RKC_FIRST, M_RK_MID) }
        else if (ma == M_RK_MID && mb == M_RKC_LAST) { SWM_FS(M_RK_MID, M_RKC_LAST) }
        else if (ma == M_RKC_FIRST && mb == M_RKC_LAST) { SWM_FS(M_RKC_FIRST, M_RKC_LAST) }
        else return ctx->fail(SWM_E_ARG, "fused stages %d+%d", ma, mb);
#undef SWM_FS
        HIPCHK(hipGetLastError());
        if (s2 != ctx->stream) {                             // join: the step goes on after both
            HIPCHK(hipEventRecord(ctx->join, s2));
            HIPCHK(hipStreamWaitEvent(ctx->stream, ctx->join, 0));
        }
        return SWM_OK;
    }

    // semi-implicit continuity of the SSPRK schemes in mode m
    static int continuity(swm_ctx *ctx, int m, Slots in, Slots pr, Slots ac, Slots o1, Slots o2, const Coef<Q> &k,
                          const Times &T) {
        if constexpr (F::mixed) {
            return ctx->fail(SWM_E_UNSUPPORTED, "SSPRK schemes with mixed formats");
        } else {
            const bool per = ctx->c.bc == SWM_PERIODIC;
            if (m <= M_RK_LAST || !slots_ok(ctx, m, in, pr, ac, o1, o2, true))
                return ctx->fail(SWM_E_STATE, "continuity mode %d: buffer slot missing", m);
            StageArgs<F> A = args(ctx, in, pr, ac, o1, o2, k, T);
            if (ctx->rec) {
                ctx->rec->add(3, m, -1, &A, sizeof A);
                return SWM_OK;
            }
            const dim3 grid(nblocks_waves(A.nstrips * A.nseg)), block(WAVE * WAVES_PER_BLOCK);
#define SWM_K(M_)                                                                                          \
    case M_:                                                                                               \
        if (per) hipLaunchKernelGGL((continuity_kernel<F, true, M_>), grid, block, 0, ctx->stream, A);     \
        else hipLaunchKernelGGL((continuity_kernel<F, false, M_>), grid, block, 0, ctx->stream, A);        \
        break;
            switch (m) {
                SWM_K(M_SSP2) SWM_K(M_SSP2_LAST) SWM_K(M_SSP3) SWM_K(M_SSP3_KN) SWM_K(M_S4) SWM_K(M_S4_3)
            default: return ctx->fail(SWM_E_ARG, "continuity mode %d", m);
            }
#undef SWM_K
            HIPCHK(hipGetLastError());
            return SWM_OK;
        }
    }

    // The first n_stages stages of the configured scheme (time_integration.jl:53-196) from
    // Prog.  On return *u0 and *u1 name the slots of the reference's u0 and u1.
    static int total_stages(const swm_config &c) {
        return c.time_scheme == SWM_RK ? c.RKo : c.time_scheme == SWM_SSPRK2 ? c.RKs : c.time_scheme == SWM_SSPRK3 ? c.ssp_s : 4;
    }

    // keep_u1: leave the last stage's u1 in a slot (a fused pair ending with the scheme's last
    // stage keeps it in registers only)
    static int stages(swm_ctx *ctx, int n_stages, int64_t t, Slots *u0, Slots *u1, bool keep_u1 = false) {
        const swm_config &c = ctx->c;
        if (n_stages > total_stages(c)) n_stages = total_stages(c);
        if (ctx->advcor_pending) {                           // advection_coriolis! on the initial state (:20-31)
            int rc = advcor(ctx, Slots{ctx->pu, ctx->pe}, true, true);
            if (rc) return rc;
            ctx->advcor_pending = false;
        }
        const Times T = forcing_times(ctx, t);
        const Slots P{ctx->pu, ctx->pe}, none{};
        Slots cur = P, z0 = P;
        Coef<Q> k{};
        int rc = SWM_OK;
        if (c.time_scheme == SWM_RK) {                                     // :55-78
            const Slots a{pick({P.uv}), pick({P.e})};
            const bool cp = c.compensated != 0;                            // increments summed apart, Kahan add
            auto mode = [&](int s) {
                return s == 1 ? (cp ? M_RKC_FIRST : M_RK_FIRST) : (s == c.RKo ? (cp ? M_RKC_LAST : M_RK_LAST) : M_RK_MID);
            };
            auto coef = [&](int s) {
                Coef<Q> kk{};
                kk.a = (Q)c.RKaDt[s - 1];
                kk.b = s < c.RKo ? (Q)c.RKbDt[s - 1] : Q(0);
                return kk;
            };
            for (int s = 1; s <= n_stages && rc == SWM_OK; ++s) {
                // the next u1: a free Tprog slot, or the other T slot with mixed formats
                const Slots n = ctx->mixed ? Slots{cur.t ? 1 - cur.uv : 0, cur.t ? 1 - cur.e : 0, true}
                                           : Slots{pick({P.uv, a.uv, cur.uv}), pick({P.e, a.e, cur.e})};
                if (s + 1 <= n_stages && fusable(ctx, mode(s), mode(s + 1)) && !(keep_u1 && s + 1 == c.RKo)) {   // stages s, s+1 in one kernel
                    const int s2 = s + 1;
                    // the first stage reads u0 only after stage 1 wrote it
                    rc = fused(ctx, mode(s), mode(s2), cur, P, s == 1 ? none : a, s2 < c.RKo ? n : none, a, coef(s),
                               coef(s2), T);
                    if (s2 < c.RKo) cur = n;
                    ++s;
                    continue;
                }
                rc = stage(ctx, mode(s), cur, P, a, s < c.RKo ? n : none, a, coef(s), T);
                if (s < c.RKo) cur = n;
            }
            z0 = a;
        } else if (c.time_scheme == SWM_SSPRK2) {                          // :82-110
            k.c = (Q)c.Dt_Ds;
            k.x = (Q)(1.0 / c.RKs);
            k.y = (Q)((double)(c.RKs - 1) / c.RKs);
            for (int s = 1; s <= n_stages && rc == SWM_OK; ++s) {
                const bool last = s == c.RKs;
                const Slots n{pick({P.uv, cur.uv}), pick({P.e, cur.e})};
                const Slots z = last ? Slots{pick({P.uv, cur.uv, n.uv}), pick({P.e, cur.e, n.e})} : none;
                const int m = last ? M_SSP2_LAST : M_SSP2;
                rc = stage(ctx, m, cur, P, none, n, z, k, T);
                if (!rc) rc = continuity(ctx, m, Slots{n.uv, cur.e}, P, none, n, z, k, T);
                cur = n;
                if (last) z0 = z;
            }
        } else if (c.time_scheme == SWM_SSPRK3) {                          // :114-156
            k.c = (Q)c.ssp_Dt_Dn;
            k.x = (Q)c.ssp_kna;
            k.y = (Q)c.ssp_knb;
            k.z = (Q)c.ssp_Dt_Dnc;
            for (int s = 2; s <= n_stages + 1 && s <= c.ssp_s + 1 && rc == SWM_OK; ++s) {
                const Slots n{pick({P.uv, cur.uv, z0.uv}), pick({P.e, cur.e, z0.e})};
                const int m = s == c.ssp_kn ? M_SSP3_KN : M_SSP3;
                rc = stage(ctx, m, cur, P, z0, n, none, k, T);
                if (!rc) rc = continuity(ctx, m, Slots{n.uv, cur.e}, P, z0, n, none, k, T);
                cur = n;
                if (s == c.ssp_mn || s == c.ssp_s + 1) z0 = n;            // copyto!(u0,u1) :150-155
            }
        } else {                                                            // 4SSPRK3 :160-195
            k.c = (Q)c.Dt_D;
            k.h = Q(0.5);
            k.x = (Q)(2.0 / 3.0);
            k.y = (Q)(1.0 / 3.0);
            for (int s = 1; s <= n_stages && s <= 4 && rc == SWM_OK; ++s) {
                const int n1 = pick({P.uv, cur.uv});
                const int n2 = s == 3 ? pick({P.uv, cur.uv, n1}) : -1;
                const int ne = pick({P.e, cur.e});
                const int m = s == 3 ? M_S4_3 : M_S4;
                rc = stage(ctx, m, cur, P, none, Slots{n1, -1}, Slots{n2, -1}, k, T);
                if (!rc) rc = continuity(ctx, m, Slots{n1, cur.e}, P, none, Slots{-1, ne}, none, k, T);
                cur = Slots{s == 3 ? n2 : n1, ne};
                if (s == 4) z0 = cur;                                       // copyto!(u0,u1) :191-193
            }
        }
        *u0 = z0;
        *u1 = cur;
        return rc;
    }

    static int dissipation(swm_ctx *ctx, Slots in, int out_uv) {
        const swm_config &c = ctx->c;
        DissArgs<F> A;
        std::memset((void *)&A, 0, sizeof A);   // padding too (compared byte for byte by ensembles)
        A.ar = ctx->aref();
        A.u0 = ctx->u[in.uv].ref(); A.v0 = ctx->v[in.uv].ref(); A.e0 = ctx->e[in.e].ref();
        A.out_u = ctx->u[out_uv].ref(); A.out_v = ctx->v[out_uv].ref();
        A.H = ctx->H.ref();
        A.cD = mkR<R>(c.cD); A.rD = mkR<R>(c.rD); A.cS = mkR<R>(c.cSmag); A.nuB = mkR<R>(c.nuB);
        A.sinv = mkR<R>(ctx->sci);
        A.dtd = (Q)c.Dt_diff; A.oma = (Q)c.one_minus_alpha;
        A.cm = c.Tcomm;
        A.nx = c.nx; A.ny = c.ny; A.P = ctx->P; A.L = ctx->Ld;
        A.nstrips = nstrips(ctx); A.nseg = nseg(ctx, ctx->Ld);
        if (ctx->rec) {
            ctx->rec->add(2, 0, 0, &A, sizeof A);
            return SWM_OK;
        }
        const dim3 grid(nblocks_waves(A.nstrips * A.nseg)), block(WAVE * WAVES_PER_BLOCK);
        const bool per = c.bc == SWM_PERIODIC, smag = c.diffusion == SWM_DIFF_SMAGORINSKY;
        const KTrace kt(ctx, KT_DISSIPATION);
#define SWM_DISS(PER_, D_, S_) hipLaunchKernelGGL((dissipation_kernel<F, PER_, D_, S_>), grid, block, 0, ctx->stream, A)
#define SWM_DISS_D(PER_, S_)                                   \
    switch (c.bottom_drag) {                                   \
    case SWM_DRAG_QUADRATIC: SWM_DISS(PER_, 0, S_); break;     \
    case SWM_DRAG_LINEAR: SWM_DISS(PER_, 1, S_); break;        \
    default: SWM_DISS(PER_, 2, S_); break;                     \
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

    static TracerArgs<F> tracer_args(const swm_ctx *ctx, int uv_slot, int src, int dst) {
        const swm_config &c = ctx->c;
        TracerArgs<F> a;
        a.u = ctx->at<Q>(ctx->u[uv_slot]);
        a.v = ctx->at<Q>(ctx->v[uv_slot]);
        a.um = ctx->at<ST>(ctx->um);
        a.vm = ctx->at<ST>(ctx->vm);
        a.sst = ctx->at<Q>(ctx->sst[src]);
        a.ssti = ctx->at<Q>(ctx->sst[dst]);
        a.hdu = mkR<R>(c.half_dtadvu); a.hdv = mkR<R>(c.half_dtadvv); a.du_ = mkR<R>(c.dtadvu); a.dv_ = mkR<R>(c.dtadvv);
        a.jSST = mkR<R>(c.jSST); a.SSTmin = mkR<R>(c.SSTmin); a.sinv = mkR<R>(ctx->sci);
        a.nx = c.nx; a.ny = c.ny; a.P = ctx->P; a.ep = ctx->ep; a.consume = c.tracer_consumption; a.cm = c.Tcomm;
        a.bad = ctx->dflags + 1;
        return a;
    }

    static int tracer_advect(swm_ctx *ctx, int uv_slot, int src, int dst) {
        const swm_config &c = ctx->c;
        const dim3 grid((c.nx + 127) / 128, c.ny), block(128);
        const bool per = c.bc == SWM_PERIODIC;
        auto k = per ? tracer_kernel<F, true> : tracer_kernel<F, false>;
        const KTrace kt(ctx, KT_TRACER);
        hipLaunchKernelGGL(k, grid, block, 0, ctx->stream, tracer_args(ctx, uv_slot, src, dst));
        HIPCHK(hipGetLastError());
        return SWM_OK;
    }

    // tracer!  tracer_advection.jl:1-44 with the post-dissipation velocities (:226-228)
    static int tracer(swm_ctx *ctx, int64_t i, int uv_slot) {
        const swm_config &c = ctx->c;
        if (!c.tracer_advection) return SWM_OK;
        const size_t n = ctx->elems;
        if ((i + c.nadvstep_half) % c.nadvstep == 0) {      // mid-point snapshot um, vm ← T(u0), T(v0)  :12-17
            snapshot_kernel<F><<<1024, 256, 0, ctx->stream>>>((ST *)ctx->um.base, (const Q *)ctx->u[uv_slot].base, n);
            snapshot_kernel<F><<<1024, 256, 0, ctx->stream>>>((ST *)ctx->vm.base, (const Q *)ctx->v[uv_slot].base, n);
            HIPCHK(hipGetLastError());
        }
        if (i % c.nadvstep == 0) {                          // :19-42
            const int src = ctx->rsst, dst = 1 - ctx->rsst;
            int rc = tracer_advect(ctx, uv_slot, src, dst);
            if (rc) return rc;
            ctx->rsst = dst;
        }
        return SWM_OK;
    }

    static int ghosts(swm_ctx *ctx, int uv_slot, int e_slot, int sst_slot) {
        const swm_config &c = ctx->c;
        const bool per = c.bc == SWM_PERIODIC;
        auto k = per ? ghost_kernel<Q, true> : ghost_kernel<Q, false>;
        Q *u = ctx->at<Q>(ctx->u[uv_slot]), *v = ctx->at<Q>(ctx->v[uv_slot]), *e = ctx->at<Q>(ctx->e[e_slot]);
        Q *s = ctx->at<Q>(ctx->sst[sst_slot]);
        hipLaunchKernelGGL(k, dim3((c.ny + 4 + 127) / 128), dim3(128), 0, ctx->stream, u, v, e, s, c.nx, c.ny, ctx->P,
                           (Q)c.one_minus_alpha, c.Tcomm, 0);
        hipLaunchKernelGGL(k, dim3((c.nx + 5 + 127) / 128), dim3(128), 0, ctx->stream, u, v, e, s, c.nx, c.ny, ctx->P,
                           (Q)c.one_minus_alpha, c.Tcomm, 1);
        HIPCHK(hipGetLastError());
        return SWM_OK;
    }

    // NaN check of Prog (u, v of slot uv, η of slot e) and sst, feedback.jl:61-72
    static int nan_check(swm_ctx *ctx, int uv_slot, int e_slot, int sst_slot, bool *found) {
        const swm_config &c = ctx->c;
        NanArgs A;
        A.f[0] = ctx->at<Q>(ctx->u[uv_slot]); A.x0[0] = -1 - ctx->ep; A.x1[0] = ctx->nux + 2 - ctx->ep; A.y0[0] = -1; A.y1[0] = c.ny + 2;
        A.f[1] = ctx->at<Q>(ctx->v[uv_slot]); A.x0[1] = -1; A.x1[1] = c.nx + 2; A.y0[1] = -1; A.y1[1] = c.ny + 1;
        A.f[2] = ctx->at<Q>(ctx->e[e_slot]); A.x0[2] = 0; A.x1[2] = c.nx + 1; A.y0[2] = 0; A.y1[2] = c.ny + 1;
        A.f[3] = ctx->at<Q>(ctx->sst[sst_slot]); A.x0[3] = 0; A.x1[3] = c.nx + 1; A.y0[3] = 1; A.y1[3] = c.ny;
        A.P = ctx->P;
        A.nf = 4;
        HIPCHK(hipMemsetAsync(ctx->dflags, 0, sizeof(int), ctx->stream));
        nan_kernel<Q><<<dim3(1, c.ny + 4, 4), 256, 0, ctx->stream>>>(A, ctx->dflags);
        HIPCHK(hipGetLastError());
        int flag = 0;
        HIPCHK(hipMemcpyAsync(&flag, ctx->dflags, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        *found = flag != 0;
        return SWM_OK;
    }

    // one pass of the loop body, time_integration.jl:45-243
    static int step(swm_ctx *ctx, int64_t i, int64_t t, bool *nan) {
        const swm_config &c = ctx->c;
        *nan = false;
        const KTrace kt(ctx, KT_STEP);
        Slots u0, u1;
        int rc = stages(ctx, 1 << 20, t, &u0, &u1);
        if (rc) return rc;
        if (adv_form(c) & ADV_KEEP) {                        // :208-211, and bottom_drag_quadratic!'s KEu, KEv
            const bool due = i % c.nstep_advcor == 0;
            if (due || (c.bottom_drag == SWM_DRAG_QUADRATIC && i % c.nstep_diff == 0))
                if ((rc = advcor(ctx, u0, due, false))) return rc;
        }
        int new_uv = u0.uv;
        if (i % c.nstep_diff == 0) {                         // :216-221
            const int out = pick({ctx->pu, u0.uv});
            rc = dissipation(ctx, u0, out);
            if (rc) return rc;
            new_uv = out;
        }
        rc = tracer(ctx, i, new_uv);                         // :226-228
        if (rc) return rc;
        if (c.nout > 0 && i % c.nout == 0) {                 // feedback.jl:136-148
            rc = nan_check(ctx, ctx->pu, ctx->pe, ctx->rsst, nan);
            if (rc) return rc;
            if (*nan) return SWM_OK;                         // keep Prog: no copy-back
        }
        ctx->pu = new_uv;                                    // copy-back (:240-242)
        ctx->pe = u0.e;
        return SWM_OK;
    }

    static int time_kernel(swm_ctx *ctx, int which, int reps, double *ms) {
        const swm_config &c = ctx->c;
        const Slots P{ctx->pu, ctx->pe}, none{};
        const Slots a{pick({P.uv}), pick({P.e})};
        const bool tp = ctx->mixed;
        const Slots s1 = tp ? Slots{0, 0, true} : Slots{pick({P.uv, a.uv}), pick({P.e, a.e})};
        const Slots s2 = tp ? Slots{1, 1, true} : Slots{pick({P.uv, a.uv, s1.uv}), pick({P.e, a.e, s1.e})};
        const int dq = pick({P.uv, a.uv});                   // dissipation output (Tprog)
        const Times T{mkR<R>(1.0), mkR<R>(1.0), mkR<R>(0.0)};
        Coef<Q> k{};
        k.a = (Q)c.RKaDt[0];
        k.b = (Q)c.RKbDt[0];
        k.c = (Q)c.Dt_D;
        auto launch = [&]() -> int {
            switch (which) {
            case 0: return stage(ctx, c.time_scheme == SWM_RK ? M_RK_MID : M_SSP2, s1, P, a, s2, a, k, T) ||
                           (c.time_scheme != SWM_RK && continuity(ctx, M_SSP2, Slots{s2.uv, s1.e}, P, none, s2, none, k, T));
            case 1: return stage(ctx, c.time_scheme == SWM_RK ? M_RK_FIRST : M_SSP2, P, P, a, s1, a, k, T);
            case 2: return stage(ctx, c.time_scheme == SWM_RK ? M_RK_LAST : M_SSP2, s1, P, a, none, a, k, T);
            case 3: return dissipation(ctx, a, dq);
            case 4: return tracer_advect(ctx, P.uv, ctx->rsst, 1 - ctx->rsst);
            case 5: return fused(ctx, M_RK_FIRST, M_RK_MID, P, P, none, s1, a, k, k, T);
            case 6: return fused(ctx, M_RK_MID, M_RK_LAST, s1, P, a, none, a, k, k, T);
            }
            return ctx->fail(SWM_E_ARG, "unknown kernel %d", which);
        };
        int rc = launch();                                   // warm-up
        if (rc) return rc;
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, ctx->stream));
        for (int r = 0; r < reps && rc == SWM_OK; ++r) rc = launch();
        HIPCHK(hipEventRecord(e1, ctx->stream));
        HIPCHK(hipEventSynchronize(e1));
        float tm = 0.0f;
        HIPCHK(hipEventElapsedTime(&tm, e0, e1));
        HIPCHK(hipEventDestroy(e0));
        HIPCHK(hipEventDestroy(e1));
        *ms = (double)tm / reps;
        return rc;
    }
};

template <class F> int run_steps(swm_ctx *ctx, int64_t nsteps, int64_t i0, int64_t t0, int64_t *done, int64_t *nan_step) {
    *done = 0;
    *nan_step = 0;
    for (int64_t k = 0; k < nsteps; ++k) {
        bool nan = false;
        int rc = Launch<F>::step(ctx, i0 + k, t0 + k * ctx->c.dtint, &nan);
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

template <class Fm> struct Adj {
    using R = typename Fm::R;
    using Q = typename Fm::Q;
    using LF = Launch<Fm>;

    static int check(swm_ctx *ctx) {
        const swm_config &c = ctx->c;
        if (c.time_scheme == SWM_SSPRK3 && c.ssp_s + 2 > 12)            // the taped step keeps s+2 slot sets of 12
            return ctx->fail(SWM_E_UNSUPPORTED, "the adjoint of the SSPRK3 scheme supports RKn = 2 or 3");
        if (c.time_scheme == SWM_SSPRK2 && (c.RKs < 2 || c.RKs > 3))   // the taped step keeps RKs+3 velocity slots of 6
            return ctx->fail(SWM_E_UNSUPPORTED, "the adjoint of the SSPRK2 scheme supports RKs = 2 or 3 stages");
        if (c.T != c.Tprog || c.Tcomm != c.Tprog || (c.T != SWM_F64 && c.T != SWM_F32))
            return ctx->fail(SWM_E_UNSUPPORTED, "the adjoint needs T = Tprog = Tcomm in {Float64, Float32}");
        if (c.scale_log2 != 0 || c.compensated)
            return ctx->fail(SWM_E_UNSUPPORTED, "the adjoint is of the unscaled, uncompensated step");
        const int af = adv_form(c);
        if (af != ADV_AH && af != ADV_SAD && af != ADV_LIN && af != (ADV_AH | ADV_KEEP) && af != (ADV_SAD | ADV_KEEP))
            return ctx->fail(SWM_E_UNSUPPORTED, "the adjoint is of the nonlinear right-hand side with the Arakawa-Hsu or "
                                                "Sadourny scheme or of the linear dynamics");
        if (c.dynamics == SWM_LINEAR && !ctx->have_fuv)
            return ctx->fail(SWM_E_STATE, "the linear dynamics need f_u, f_v (swm_set_static_linear)");
        if (!ctx->have_static || !ctx->have_prog) return ctx->fail(SWM_E_STATE, "set_static and upload_prog first");
        if (ctx->ens_refs) return ctx->fail(SWM_E_STATE, "the context belongs to an ensemble; destroy it first");
        return SWM_OK;
    }

    // slot sets the taped step keeps: Prog, u0, the dissipation output and the stage inputs
    // (classic RK: s_2..s_RKo; SSPRK2: w_2..w_RKs+1; 4SSPRK3: w_2, w_3, (u'_3, v'_3), w_4, so 7;
    // SSPRK3: w_2..w_s+1 with u0 = w_s+1, so s+2)
    static int adj_slots(const swm_config &c) {
        return c.time_scheme == SWM_4SSPRK3 ? 7 : c.time_scheme == SWM_SSPRK3 ? std::max(6, c.ssp_s + 2) : 6;
    }

    // carve the adjoint's fields: the taped step's slot sets beyond the forward's four and the
    // adjoint fields (28), zeroed
    static int init(swm_ctx *ctx) {
        if (ctx->adj) return SWM_OK;
        int rc = check(ctx);
        if (rc) return rc;
        const int nsl = adj_slots(ctx->c);                // slot sets of the taped step
        const int nf2 = 3 * (nsl - 4) + 28;
        const size_t n1d = (size_t)ctx->c.ny + 2 * OY + 16;
        if ((rc = grow_arena(ctx, (size_t)nf2 * (ctx->elems * ctx->es + 256) + n1d * ctx->es + 256))) return rc;
        AdjState *st = new AdjState();
        const int es = ctx->es;
        for (int s = 4; s < nsl; ++s)
            if ((rc = alloc_field(ctx, ctx->u[s], 0, true, es)) || (rc = alloc_field(ctx, ctx->v[s], 0, true, es)) ||
                (rc = alloc_field(ctx, ctx->e[s], 0, true, es))) {
                delete st;
                return rc;
            }
        auto one = [&](uint32_t &o, bool two_d) {
            Field f;
            if (!rc) rc = alloc_field(ctx, f, two_d ? 0 : n1d, two_d, es);
            o = f.origin;
        };
        auto uve = [&](adj::UVE &x) { one(x.u, true); one(x.v, true); one(x.e, true); };
        adj::AdjFields &F = st->F;
        uve(F.L); uve(F.Z); uve(F.Ma); uve(F.Mb); uve(F.A);
        one(F.Lsst, true); one(F.Lsst2, true); one(F.Lum, true); one(F.Lvm, true);
        one(F.tr0, true); one(F.tr1, true);
        one(F.tg0, true); one(F.tg1, true);
        one(F.lfx, true);
        one(F.lqend, false);
        one(F.qk, true); one(F.lqk, true); one(F.lkk, true);
        one(F.zero, true);
        if (rc) {
            delete st;
            return rc;
        }
        F.H = ctx->H.ref();
        F.fq = ctx->fq.ref();
        F.fq_2d = ctx->fq_sx;
        F.fu = ctx->f_u.ref(); F.fu_2d = ctx->fu_sx;
        F.fv = ctx->f_v.ref(); F.fv_2d = ctx->fv_sx;
        F.qend = ctx->qend.ref();
        st->F0 = F;
        // rows per wave: about four resident waves per wave slot, at least 6 (a sweep of 4 .. 64 at
        // 1024² and 4096² finds the result within 1 % of the best at 4096² and 6 the best at 1024²)
        st->La = LF::rows_per_wave(ctx, reinterpret_cast<const void *>(&stage_adj_kernel<R, false, true, 0, false>),
                                   adj::nstrips_for<adj::RT_HALO>(ctx->c.nx + 2), 4, 6);
        if (const char *e = getenv("SWM_ROWS_PER_WAVE_ADJ")) {
            const int l = atoi(e);
            if (l >= 4 && l <= 4096) st->La = l;
        }
        ctx->adj = st;
        ctx->nslots = nsl;
        return SWM_OK;
    }

    // the constants of the reverse of a step at model time t
    static adj::AdjConsts<R> consts(const swm_ctx *ctx, int64_t t) {
        const swm_config &c = ctx->c;
        adj::AdjConsts<R> k{};
        k.fxt = (R)(c.seasonal_wind_x ? round_host(std::sin(c.omegaFx * (double)t), c.T) : 1.0);
        k.fxg = ctx->adj_fx ? 1 : 0;
        for (int j = 0; j < 4; ++j) {
            k.a[j] = (R)c.RKaDt[j];
            k.b[j] = j < 3 ? (R)c.RKbDt[j] : R(0);
        }
        k.sad = c.adv_scheme == SWM_ADV_SADOURNY;
        k.lin = c.dynamics == SWM_LINEAR;
        k.nadvcor = (adv_form(c) & ADV_KEEP) ? (int)c.nstep_advcor : 0;
        if (c.time_scheme == SWM_SSPRK2) {               // the forward's coefficients, in the same format
            k.ssp = c.RKs;
            k.cs = (R)c.Dt_Ds;
            k.sa = (R)(1.0 / c.RKs);
            k.sb = (R)((double)(c.RKs - 1) / c.RKs);
        } else if (c.time_scheme == SWM_SSPRK3) {        // stages counted from 1 (the reference's rki = 2 .. s+1)
            k.s3 = c.ssp_s;
            k.kn = c.ssp_kn - 1;
            k.mn = c.ssp_mn - 1;
            k.cs = (R)c.ssp_Dt_Dn;
            k.cs2 = (R)c.ssp_Dt_Dnc;
            k.sa = (R)c.ssp_kna;
            k.sb = (R)c.ssp_knb;
        } else if (c.time_scheme == SWM_4SSPRK3) {       // u' = ½(u + (u + Δt_Δ·du)): ½·Δt_Δ on du (exact)
            k.s4 = 1;
            k.cs = (R)(0.5 * (double)(R)c.Dt_D);
            k.sa = (R)(2.0 / 3.0);
            k.sb = (R)(1.0 / 3.0);
        }
        k.g = (R)c.g; k.gamma = (R)c.gamma; k.k12 = (R)round_host(k.sad ? 0.5 : 1.0 / 12.0, c.T); k.oma = (R)c.one_minus_alpha;
        k.cD = (R)c.cD; k.rD = (R)c.rD; k.cS = (R)c.cSmag; k.nuB = (R)c.nuB; k.dtd = (R)c.Dt_diff;
        k.hdu = (R)c.half_dtadvu; k.hdv = (R)c.half_dtadvv; k.du = (R)c.dtadvu; k.dv = (R)c.dtadvv; k.jSST = (R)c.jSST;
        k.RKo = c.RKo; k.cont = c.continuity; k.drag = c.bottom_drag; k.smag = c.diffusion == SWM_DIFF_SMAGORINSKY;
        k.per = c.bc == SWM_PERIODIC; k.consume = c.tracer_consumption; k.tracer = c.tracer_advection;
        k.nadvstep = (int)c.nadvstep; k.nadvstep_half = (int)c.nadvstep_half; k.nstep_diff = c.nstep_diff;
        k.nx = c.nx; k.ny = c.ny; k.P = ctx->P; k.ep = ctx->ep;
        return k;
    }

    // ---- kernel launcher of adj::reverse_step
    struct K {
        swm_ctx *ctx;
        AdjState *st;
        int fail_launch() {
            hipError_t e = hipGetLastError();
            return e == hipSuccess ? SWM_OK : ctx->fail(SWM_E_HIP, "adjoint kernel launch: %s", hipGetErrorString(e));
        }
        template <bool PER, bool NEXT, int NX, bool PRE, int SCH> void launch_stage(const adj::StageAdjArgs<R> &S) {
            hipLaunchKernelGGL((stage_adj_kernel<R, PER, NEXT, NX, PRE, SCH>), dim3(nblocks_waves(S.nstrips * S.nseg)),
                               dim3(WAVE * WAVES_PER_BLOCK), 0, ctx->stream, ctx->aref(), S);
        }
        template <bool PER, int SCH> void stage_p(const adj::StageAdjArgs<R> &S, int next, int nx, int pre) {
            if (!next) launch_stage<PER, false, 0, false, SCH>(S);
            else if (!pre) launch_stage<PER, true, 0, false, SCH>(S);
            else if (nx == 0) launch_stage<PER, true, 0, true, SCH>(S);
            else if (nx == 1) launch_stage<PER, true, 1, true, SCH>(S);
            else launch_stage<PER, true, 2, true, SCH>(S);
        }
        template <bool PER> void stage_s(const adj::StageAdjArgs<R> &S, int next, int nx, int pre, int sch) {
            if (sch == adj::ADJ_CONT) launch_stage<PER, false, 0, false, adj::ADJ_CONT>(S);   // (no successor)
            else if (sch == adj::ADJ_LIN) stage_p<PER, adj::ADJ_LIN>(S, next, nx, pre);
            else if (sch == adj::ADJ_SAD) stage_p<PER, adj::ADJ_SAD>(S, next, nx, pre);
            else if (sch == (adj::ADJ_AH | adj::ADJ_KEEP)) stage_p<PER, adj::ADJ_AH | adj::ADJ_KEEP>(S, next, nx, pre);
            else if (sch == (adj::ADJ_SAD | adj::ADJ_KEEP)) stage_p<PER, adj::ADJ_SAD | adj::ADJ_KEEP>(S, next, nx, pre);
            else stage_p<PER, adj::ADJ_AH>(S, next, nx, pre);
        }
        int advcor_adj(const adj::AdvcorAdjArgs &X) {
            hipLaunchKernelGGL((advcor_adj_kernel<R>), dim3((X.nx + 5 + 127) / 128, X.ny + 4), dim3(128), 0, ctx->stream,
                               ctx->aref(), X, ctx->P);
            return fail_launch();
        }
        int lqend_keep(const adj::StageAdjArgs<R> &S) {
            hipLaunchKernelGGL((lqend_keep_kernel<R>), dim3((S.ny + 1 + 255) / 256), dim3(256), 0, ctx->stream, ctx->aref(), S);
            return fail_launch();
        }
        int lincomb(uint32_t d, uint32_t x, R ax, uint32_t y, R ay, uint32_t z, R az) {   // 2-D fields of the arena
            char *b = (char *)ctx->arena;
            hipLaunchKernelGGL((lincomb_kernel<R>), dim3(1024), dim3(256), 0, ctx->stream, (R *)(b + d), (const R *)(b + x), ax,
                               y == adj::NOF ? nullptr : (const R *)(b + y), ay, z == adj::NOF ? nullptr : (const R *)(b + z), az,
                               ctx->elems);
            return fail_launch();
        }
        int zero(uint32_t o) {             // a 2-D field of the arena
            hipLaunchKernelGGL(zero_words_kernel, dim3(1024), dim3(256), 0, ctx->stream,
                               (uint32_t *)((char *)ctx->arena + o), ctx->elems * ctx->es / 4);
            return fail_launch();
        }
        int stage(adj::StageAdjArgs<R> S, int per, int next, int nx, int pre, int sch) {
            S.L = st->La;
            S.nstrips = adj::nstrips_for<adj::RT_HALO>(S.nx + 2);
            S.nseg = (adj::adj_y1(S.ny) - adj::ADJ_Y0 + S.L - 1) / S.L;
            if (per) stage_s<true>(S, next, nx, pre, sch);
            else stage_s<false>(S, next, nx, pre, sch);
            return fail_launch();
        }
        int qend(const adj::StageAdjArgs<R> &S) {
            hipLaunchKernelGGL((qend_fwd_kernel<R>), dim3((S.ny + 1 + 255) / 256), dim3(256), 0, ctx->stream, ctx->aref(), S);
            return fail_launch();
        }
        int qend_adj(const adj::StageAdjArgs<R> &S, const adj::UVE &mu) {
            hipLaunchKernelGGL((qend_adj_kernel<R>), dim3((S.ny + 2 + 255) / 256), dim3(256), 0, ctx->stream, ctx->aref(), S,
                               mu);
            return fail_launch();
        }
        template <bool PER, int DRAG, bool SMAG> void launch_diss(const adj::DissAdjArgs<R> &D) {
            hipLaunchKernelGGL((diss_adj_kernel<R, PER, DRAG, SMAG>), dim3(nblocks_waves(D.nstrips * D.nseg)),
                               dim3(WAVE * WAVES_PER_BLOCK), 0, ctx->stream, ctx->aref(), D);
        }
        template <bool PER> void diss_p(const adj::DissAdjArgs<R> &D, int drag, int smag) {
            if (drag == 0) { if (smag) launch_diss<PER, 0, true>(D); else launch_diss<PER, 0, false>(D); }
            else if (drag == 1) { if (smag) launch_diss<PER, 1, true>(D); else launch_diss<PER, 1, false>(D); }
            else { if (smag) launch_diss<PER, 2, true>(D); else launch_diss<PER, 2, false>(D); }
        }
        int diss(adj::DissAdjArgs<R> D, int per, int drag, int smag) {
            D.L = st->La;
            D.nstrips = adj::nstrips_for<adj::DT_HALO>(D.nx + 2);
            D.nseg = (adj::adj_y1(D.ny) - adj::ADJ_Y0 + D.L - 1) / D.L;
            if (per) diss_p<true>(D, drag, smag);
            else diss_p<false>(D, drag, smag);
            return fail_launch();
        }
        int tracer(const adj::TracerAdjArgs<R> &T) {
            hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(64), 0, ctx->stream, (uint32_t *)st->dmax, (size_t)2);
            // the first pass: 256 columns per block, rows per thread so that about 8192 blocks run
            const int bx = (T.nx + 255) / 256, rows = std::max(1, (bx * T.ny + 8191) / 8192);
            // the gather of (a): corners X = -2 .. nx+3, Y = -1 .. ny+3 on tiles of GT_X x gt_y<R>
            // (SWM_TRACER_GATHER_GLOBAL=1: every corner scans in global memory, as a check)
            const int tiled = getenv("SWM_TRACER_GATHER_GLOBAL") ? 0 : 1;
            const dim3 g1(bx, (T.ny + rows - 1) / rows), ga((T.nx + 6 + GT_X - 1) / GT_X, (T.ny + 5 + gt_y<R> - 1) / gt_y<R>),
                g2((T.nx + 5 + 127) / 128, T.ny + 4);
            if (T.ep) {
                hipLaunchKernelGGL((tracer_src_kernel<R, true>), g1, dim3(256), 0, ctx->stream, ctx->aref(), T, rows, st->dmax);
                hipLaunchKernelGGL((tracer_gather_uv_tiled_kernel<R, true>), ga, dim3(256), 0, ctx->stream, ctx->aref(), T, st->dmax,
                                   tiled);
                hipLaunchKernelGGL((tracer_gather_kernel<R, true>), g2, dim3(128), 0, ctx->stream, ctx->aref(), T);
            } else {
                hipLaunchKernelGGL((tracer_src_kernel<R, false>), g1, dim3(256), 0, ctx->stream, ctx->aref(), T, rows, st->dmax);
                hipLaunchKernelGGL((tracer_gather_uv_tiled_kernel<R, false>), ga, dim3(256), 0, ctx->stream, ctx->aref(), T, st->dmax,
                                   tiled);
                hipLaunchKernelGGL((tracer_gather_kernel<R, false>), g2, dim3(128), 0, ctx->stream, ctx->aref(), T);
            }
            return fail_launch();
        }
        int snapshot(uint32_t lu, uint32_t lum, uint32_t lv, uint32_t lvm) {
            char *b = (char *)ctx->arena;
            hipLaunchKernelGGL((snapshot_adj_kernel<R>), dim3(1024), dim3(256), 0, ctx->stream, (R *)(b + lu), (R *)(b + lum),
                               (R *)(b + lv), (R *)(b + lvm), ctx->elems);
            return fail_launch();
        }
    };

    // ---- snapshots of full states (u, v, η, sst, um, vm)
    // with nstep_advcor > 0 also the kept q and KEu+KEv and whether they are still to be formed
    static int snap_fields(const swm_ctx *ctx) { return (adv_form(ctx->c) & ADV_KEEP) ? 8 : 6; }
    // `count` snapshots in one device block; a block of the same layout is reused
    static int snap_alloc(swm_ctx *ctx, int count) {
        AdjState *st = ctx->adj;
        const size_t bytes = (size_t)snap_fields(ctx) * ctx->elems * ctx->es;
        st->snap_pending.assign(count, 0);
        if (!(st->snap_block && (int)st->snap.size() == count && st->snap_bytes == bytes && st->dmax)) {
            st->release();
            st->snap_bytes = bytes;
            void *p = nullptr;
            if (hipMalloc(&p, (size_t)count * bytes) != hipSuccess) {
                (void)hipGetLastError();
                return ctx->fail(SWM_E_HIP, "hipMalloc of %d snapshots (%zu bytes) failed", count, (size_t)count * bytes);
            }
            st->snap_block = (char *)p;
            for (int k = 0; k < count; ++k) st->snap.push_back(st->snap_block + (size_t)k * bytes);
            if (hipMalloc(&st->dmax, 2 * sizeof(int)) != hipSuccess) return ctx->fail(SWM_E_HIP, "hipMalloc");
        }
        st->at.assign(count, -1);
        return SWM_OK;
    }

    // snapshots < 0: as many as half of the device's free memory holds (beside the ones a
    // schedule always has), at most nsteps - 1: every state kept, nothing recomputed
    static int auto_snapshots(swm_ctx *ctx, int64_t n) {
        size_t freeb = 0, total = 0;
        AdjState *st = ctx->adj;
        if (st->snap_block) freeb += st->snap.size() * st->snap_bytes;     // (reused or freed by snap_alloc)
        size_t f = 0;
        if (hipMemGetInfo(&f, &total) != hipSuccess) f = 0;
        freeb += f;
        const size_t bytes = (size_t)snap_fields(ctx) * ctx->elems * ctx->es;
        const int64_t fit = (int64_t)(freeb / 2 / bytes) - 2;
        return (int)std::max<int64_t>(0, std::min<int64_t>({fit, n - 1, (int64_t)INT32_MAX - 2}));
    }
    static int copy_state(swm_ctx *ctx, int k, bool save) {
        AdjState *st = ctx->adj;
        const size_t fb = ctx->elems * ctx->es;
        const Field *f[8] = {&ctx->u[ctx->pu], &ctx->v[ctx->pu], &ctx->e[ctx->pe], &ctx->sst[ctx->rsst], &ctx->um, &ctx->vm,
                             &ctx->qs, &ctx->ks};
        const int nf = snap_fields(ctx);
        // one launch for every field when all are 16-byte aligned, else one copy each
        CopySet S{};
        S.n = nf;
        S.n16 = fb / 16;
        bool vec = fb % 16 == 0;
        for (int j = 0; j < nf; ++j) {
            char *s = st->snap[k] + j * fb;
            char *src = save ? (char *)f[j]->base : s, *dst = save ? s : (char *)f[j]->base;
            vec = vec && ((uintptr_t)src | (uintptr_t)dst) % 16 == 0;
            S.src[j] = (const uint4 *)src;
            S.dst[j] = (uint4 *)dst;
        }
        if (vec) {
            const int blocks = (int)std::min<size_t>(2048, (S.n16 + 255) / 256);
            hipLaunchKernelGGL(copy_fields_kernel, dim3(std::max(1, blocks)), dim3(256), 0, ctx->stream, S);
            HIPCHK(hipGetLastError());
        } else
            for (int j = 0; j < nf; ++j)
                HIPCHK(hipMemcpyAsync(S.dst[j], S.src[j], fb, hipMemcpyDeviceToDevice, ctx->stream));
        if (save) st->snap_pending[k] = ctx->advcor_pending ? 1 : 0;
        else ctx->advcor_pending = st->snap_pending[k] != 0;
        return SWM_OK;
    }

    // one forward step from Prog at step i (0-based k = i - i0)
    static int advance(swm_ctx *ctx, int64_t k) {
        AdjState *st = ctx->adj;
        bool nan = false;
        const int64_t i = st->i0 + k;
        int rc = LF::step(ctx, i, st->t0 + k * ctx->c.dtint, &nan);
        if (rc) return rc;
        if (nan) return ctx->fail(SWM_E_STATE, "NaN in the state at step %lld: no gradient", (long long)i);
        st->fwd_steps += 1;
        return SWM_OK;
    }

    // The taped step k: step i = i0 + k from Prog without copy-back, every intermediate kept
    // in its own slot (the stage inputs s_1..s_RKo, or w_1..w_RKs+1 under SSPRK2, or under
    // 4SSPRK3 w_1..w_4 and each stage's velocities after its momentum update; u0 before /
    // after the dissipation).
    static int tape_step(swm_ctx *ctx, int64_t k, adj::AdjFields &F) {
        AdjState *st = ctx->adj;
        const swm_config &c = ctx->c;
        const int64_t i = st->i0 + k, t = st->t0 + k * c.dtint;
        const typename LF::Times T = LF::forcing_times(ctx, t);
        int uv[12], ee[12], nu = 0, ne = 0;
        uv[nu++] = ctx->pu;
        ee[ne++] = ctx->pe;
        for (int s = 0; s < ctx->nslots; ++s) {
            if (s != ctx->pu) uv[nu++] = s;
            if (s != ctx->pe) ee[ne++] = s;
        }
        // slots: [0] Prog = s_1, [1] u0, [2..RKo] s_2..s_RKo (SSPRK2: [2..RKs+1] w_2..w_RKs+1),
        // [5] the dissipation output; 4SSPRK3: see below
        const bool s4 = c.time_scheme == SWM_4SSPRK3, s3 = c.time_scheme == SWM_SSPRK3;
        auto sk = [&](int q) { return q == 1 ? Slots{uv[0], ee[0]} : Slots{uv[q], ee[q]}; };   // input of stage q (1-based)
        // SSPRK3: w_q in sk(q), q = 1 .. s+1 (slots 0, 2 .. s+1), u0 = w_s+1, the dissipation output in uv[1]
        const Slots P{uv[0], ee[0]}, a = s4 ? Slots{uv[5], ee[4]} : s3 ? sk(c.ssp_s + 1) : Slots{uv[1], ee[1]}, none{};
        int rc = SWM_OK;
        auto coef = [&](int q) {
            Coef<Q> kk{};
            kk.a = (Q)c.RKaDt[q - 1];
            kk.b = q < c.RKo ? (Q)c.RKbDt[q - 1] : Q(0);
            return kk;
        };
        auto mode = [&](int q) { return q == 1 ? M_RK_FIRST : (q == c.RKo ? M_RK_LAST : M_RK_MID); };
        const bool keep = (adv_form(c) & ADV_KEEP) != 0;
        if (keep) {
            // q and KEu+KEv from the step's input when still to be formed (:20-31); the q the
            // stages read is taped
            if (ctx->advcor_pending) {
                if ((rc = LF::advcor(ctx, P, true, true))) return rc;
                ctx->advcor_pending = false;
            }
            HIPCHK(hipMemcpyAsync((char *)ctx->arena + F.qk, ctx->qs.base, ctx->elems * ctx->qs.es, hipMemcpyDeviceToDevice,
                                  ctx->stream));
        }
        const bool ssp = c.time_scheme == SWM_SSPRK2;
        if (ssp) {                                           // :82-110; w_q in sk(q), q = 1 .. RKs+1
            Coef<Q> kk{};
            kk.c = (Q)c.Dt_Ds;
            kk.x = (Q)(1.0 / c.RKs);
            kk.y = (Q)((double)(c.RKs - 1) / c.RKs);
            for (int q = 1; q <= c.RKs && rc == SWM_OK; ++q) {
                const bool last = q == c.RKs;
                const int m = last ? M_SSP2_LAST : M_SSP2;
                const Slots n = sk(q + 1);
                rc = LF::stage(ctx, m, sk(q), P, none, n, last ? a : none, kk, T);
                if (!rc) rc = LF::continuity(ctx, m, Slots{n.uv, sk(q).e}, P, none, n, last ? a : none, kk, T);
            }
        }
        if (s4) {                                            // :160-195
            // w_1 = P; u'_1 = w_2 in uv[1], η_2 in ee[1]; u'_2 = w_3 in uv[2], η_3 in ee[2];
            // u'_3 in uv[3], w_4 in (uv[4], ee[3]); u0 = (u'_4, η'_4) in (uv[5], ee[4])
            Coef<Q> kk{};
            kk.c = (Q)c.Dt_D;
            kk.h = Q(0.5);
            kk.x = (Q)(2.0 / 3.0);
            kk.y = (Q)(1.0 / 3.0);
            const Slots w[5] = {P, Slots{uv[1], ee[1]}, Slots{uv[2], ee[2]}, Slots{uv[4], ee[3]}, a};
            const int up[4] = {uv[1], uv[2], uv[3], uv[5]};
            for (int q = 1; q <= 4 && rc == SWM_OK; ++q) {
                const int m = q == 3 ? M_S4_3 : M_S4;
                rc = LF::stage(ctx, m, w[q - 1], P, none, Slots{up[q - 1], -1}, Slots{q == 3 ? uv[4] : -1, -1}, kk, T);
                if (!rc) rc = LF::continuity(ctx, m, Slots{up[q - 1], w[q - 1].e}, P, none, Slots{-1, w[q].e}, none, kk, T);
                F.s[q - 1] = adj::UVE{ctx->u[w[q - 1].uv].ref(), ctx->v[w[q - 1].uv].ref(), ctx->e[w[q - 1].e].ref()};
                F.p[q - 1] = adj::UVE{ctx->u[up[q - 1]].ref(), ctx->v[up[q - 1]].ref(), 0u};
            }
        }
        if (s3) {                                            // :114-156, stage q = rki - 1
            Coef<Q> kk{};
            kk.c = (Q)c.ssp_Dt_Dn;
            kk.x = (Q)c.ssp_kna;
            kk.y = (Q)c.ssp_knb;
            kk.z = (Q)c.ssp_Dt_Dnc;
            const Slots X = sk(std::max(1, c.ssp_mn));      // the state kept after stage mn (the input when mn = 1)
            for (int q = 1; q <= c.ssp_s && rc == SWM_OK; ++q) {
                const int m = q + 1 == c.ssp_kn ? M_SSP3_KN : M_SSP3;
                rc = LF::stage(ctx, m, sk(q), P, X, sk(q + 1), none, kk, T);
                if (!rc) rc = LF::continuity(ctx, m, Slots{sk(q + 1).uv, sk(q).e}, P, X, sk(q + 1), none, kk, T);
            }
        }
        for (int q = 1; !ssp && !s4 && !s3 && q <= c.RKo && rc == SWM_OK; ++q) {
            // stages q, q+1 as one fused pair that also keeps the input of stage q+1
            if (q + 1 <= c.RKo && LF::can_tape(ctx, mode(q), mode(q + 1))) {
                rc = LF::fused_tape(ctx, mode(q), mode(q + 1), sk(q), P, q == 1 ? none : a, q + 1 < c.RKo ? sk(q + 2) : none,
                                    a, sk(q + 1), coef(q), coef(q + 1), T);
                ++q;
                continue;
            }
            rc = LF::stage(ctx, mode(q), sk(q), P, a, q < c.RKo ? sk(q + 1) : none, a, coef(q), T);
        }
        if (rc) return rc;
        if (keep) {                                          // :208-211, and the quadratic drag's KEu, KEv
            const bool due = i % c.nstep_advcor == 0;
            if (due || (c.bottom_drag == SWM_DRAG_QUADRATIC && i % c.nstep_diff == 0))
                if ((rc = LF::advcor(ctx, a, due, false))) return rc;
        }
        // The reverse step reads the state after the dissipation, and what the tracer leaves,
        // only through the adjoint of the tracer's advection: a step without it ends here
        // (the mid-point snapshot of um, vm serves later forward steps, which the sweep
        // recomputes from stored states)
        const bool adv = c.tracer_advection && i % c.nadvstep == 0;
        int post = a.uv;
        if (adv && i % c.nstep_diff == 0) {
            post = s4 ? uv[6] : s3 ? uv[1] : uv[5];
            if ((rc = LF::dissipation(ctx, a, post))) return rc;
        }
        const int sst_old = ctx->rsst;
        if (adv && (rc = LF::tracer(ctx, i, post))) return rc;
        for (int q = 1; !s4 && q <= (ssp ? c.RKs + 1 : s3 ? c.ssp_s + 1 : c.RKo); ++q) {
            const Slots s = sk(q);
            F.s[q - 1] = adj::UVE{ctx->u[s.uv].ref(), ctx->v[s.uv].ref(), ctx->e[s.e].ref()};
            if ((ssp || s3) && q >= 2) F.p[q - 2] = adj::UVE{F.s[q - 1].u, F.s[q - 1].v, 0u};    // u'_k = u_{k+1}
        }
        F.pre = adj::UVE{ctx->u[a.uv].ref(), ctx->v[a.uv].ref(), ctx->e[a.e].ref()};
        F.post_u = ctx->u[post].ref();
        F.post_v = ctx->v[post].ref();
        F.um = ctx->um.ref();
        F.vm = ctx->vm.ref();
        F.sst = ctx->sst[sst_old].ref();
        st->fwd_steps += 1;
        return SWM_OK;
    }

    static int reverse_one(swm_ctx *ctx, int64_t k) {
        AdjState *st = ctx->adj;
        int rc = tape_step(ctx, k, st->F);
        if (rc) return rc;
        K kl{ctx, st};
        return adj::reverse_step<R>(kl, st->F, consts(ctx, st->t0 + k * ctx->c.dtint), (long)(st->i0 + k));
    }

    // reverse steps [a, b) with the state at a in snapshot slot sa and s free snapshots
    static int reverse(swm_ctx *ctx, int64_t a, int64_t b, int sa, int s) {
        AdjState *st = ctx->adj;
        int rc;
        auto restore_advance = [&](int64_t to) {
            int r = copy_state(ctx, sa, false);
            for (int64_t q = a; q < to && !r; ++q) r = advance(ctx, q);
            return r;
        };
        if (b - a == 1 || s == 0) {
            for (int64_t k = b - 1; k >= a; --k) {
                if ((rc = restore_advance(k)) || (rc = reverse_one(ctx, k))) return rc;
            }
            return SWM_OK;
        }
        const int64_t m = revolve_split(a, b, s);
        int sm = -1;
        for (size_t q = 1; q + 1 < st->snap.size(); ++q)
            if (st->at[q] == m) sm = (int)q;
        if (sm < 0) {                                            // not stored by the forward sweep
            for (size_t q = 1; q + 1 < st->snap.size() && sm < 0; ++q)
                if (st->at[q] < 0) sm = (int)q;
            if (sm < 0) return ctx->fail(SWM_E_STATE, "revolve: no free snapshot");
            if ((rc = restore_advance(m)) || (rc = copy_state(ctx, sm, true))) return rc;
            st->at[sm] = m;
        }
        if ((rc = reverse(ctx, m, b, sm, s - 1))) return rc;
        st->at[sm] = -1;
        return reverse(ctx, a, m, sa, s);
    }

    static int forward(swm_ctx *ctx, int64_t n, int64_t i0, int64_t t0, int s) {
        int rc = init(ctx);
        if (rc) return rc;
        AdjState *st = ctx->adj;
        if (s < 0) s = auto_snapshots(ctx, n);
        if ((rc = snap_alloc(ctx, s + 2))) return rc;
        st->i0 = i0; st->t0 = t0; st->n = n; st->nsnap = s; st->recorded = false; st->fwd_steps = 0;
        if ((rc = copy_state(ctx, 0, true))) return rc;
        st->keep_init = ctx->advcor_pending;
        // the snapshots the schedule takes first: the chain of splits toward the last step
        std::vector<int64_t> chain;
        for (int64_t a = 0, sf = s; n - a > 1 && sf > 0; --sf) {
            a = revolve_split(a, n, (int)sf);
            chain.push_back(a);
        }
        size_t next = 0;
        for (int64_t k = 0; k < n; ++k) {
            if (next < chain.size() && chain[next] == k) {
                if ((rc = copy_state(ctx, (int)next + 1, true))) return rc;
                st->at[next + 1] = k;
                ++next;
            }
            if ((rc = advance(ctx, k))) return rc;
        }
        if ((rc = copy_state(ctx, s + 1, true))) return rc;
        HIPCHK(hipStreamSynchronize(ctx->stream));
        st->recorded = true;
        return SWM_OK;
    }

    static int zero_adjoint(swm_ctx *ctx) {
        AdjState *st = ctx->adj;
        const size_t from = std::min<size_t>(st->F0.L.u, ctx->u[4].origin);
        const size_t words = (ctx->arena_used - from) / 4;
        hipLaunchKernelGGL(zero_words_kernel, dim3(2048), dim3(256), 0, ctx->stream, (uint32_t *)((char *)ctx->arena + from),
                           words);
        HIPCHK(hipGetLastError());
        return SWM_OK;
    }

    static int pullback(swm_ctx *ctx, const void *yu, const void *yv, const void *ye, const void *ys, void *gu, void *gv,
                        void *ge, void *gs) {
        AdjState *st = ctx->adj;
        const swm_config &c = ctx->c;
        int rc;
        st->F = st->F0;
        st->fwd_steps = 0;
        st->fx_valid = false;
        if ((rc = zero_adjoint(ctx))) return rc;
        Field tmp;
        auto put = [&](uint32_t o, const void *h, int m, int nn, int x0, int y0) {
            tmp.base = (char *)ctx->arena + o;
            tmp.origin = o;
            tmp.es = ctx->es;
            tmp.p0 = (char *)tmp.base + ((size_t)OX + (size_t)OY * ctx->P) * ctx->es;
            return put2d(ctx, tmp, h, m, nn, x0, y0);
        };
        const int mu = ctx->nux + 4, nuu = ctx->nuy + 4, mv = ctx->nvx + 4, nv = ctx->nvy + 4;
        if ((rc = put(st->F.L.u, yu, mu, nuu, -1 - ctx->ep, -1)) || (rc = put(st->F.L.v, yv, mv, nv, -1, -1)) ||
            (rc = put(st->F.L.e, ye, c.nx + 2, c.ny + 2, 0, 0)) || (rc = put(st->F.Lsst, ys, c.nx + 2, c.ny, 0, 1)))
            return rc;
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, ctx->stream));
        rc = reverse(ctx, 0, st->n, 0, st->nsnap);
        HIPCHK(hipEventRecord(e1, ctx->stream));
        if (rc) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return rc;
        }
        HIPCHK(hipEventSynchronize(e1));
        float tm = 0.0f;
        HIPCHK(hipEventElapsedTime(&tm, e0, e1));
        st->sweep_ms = tm;
        HIPCHK(hipEventDestroy(e0));
        HIPCHK(hipEventDestroy(e1));
        if (st->keep_init) {
            // q and KEu+KEv were formed from the initial state: their cotangents, gathered from
            // the steps before any later forming, pass to it (time_integration.jl:20-31)
            if ((rc = copy_state(ctx, 0, false))) return rc;
            adj::AdvcorAdjArgs X{};
            X.s = adj::UVE{ctx->u[ctx->pu].ref(), ctx->v[ctx->pu].ref(), ctx->e[ctx->pe].ref()};
            X.H = st->F.H; X.fq = st->F.fq; X.fq_2d = st->F.fq_2d; X.lqk = st->F.lqk; X.lkk = st->F.lkk;
            X.out = st->F.L; X.wq = 1; X.nx = c.nx; X.ny = c.ny;
            K kl{ctx, st};
            if ((rc = kl.advcor_adj(X))) return rc;
        }
        // the gradient: ghost copies of the uploaded state folded in (swm_upload_prog applies them)
        adj::FoldOutArgs O{st->F.L, st->F.Z, st->F.Lsst, st->F.Lsst2, c.nx, c.ny, c.bc == SWM_PERIODIC};
        hipLaunchKernelGGL((fold_out_kernel<R>), dim3((c.nx + 5 + 127) / 128, c.ny + 4), dim3(128), 0, ctx->stream,
                           ctx->aref(), O, (R)c.one_minus_alpha, ctx->P);
        HIPCHK(hipGetLastError());
        auto get = [&](uint32_t o, void *h, int m, int nn, int x0, int y0) {
            tmp.base = (char *)ctx->arena + o;
            tmp.origin = o;
            tmp.p0 = (char *)tmp.base + ((size_t)OX + (size_t)OY * ctx->P) * ctx->es;
            return get2d(ctx, tmp, h, m, nn, x0, y0);
        };
        if ((gu && (rc = get(st->F.Z.u, gu, mu, nuu, -1 - ctx->ep, -1))) || (gv && (rc = get(st->F.Z.v, gv, mv, nv, -1, -1))) ||
            (ge && (rc = get(st->F.Z.e, ge, c.nx + 2, c.ny + 2, 0, 0))) || (gs && (rc = get(st->F.Lsst2, gs, c.nx + 2, c.ny, 0, 1))))
            return rc;
        // leave the context at the final state, as after swm_adj_forward
        if ((rc = copy_state(ctx, st->nsnap + 1, false))) return rc;
        HIPCHK(hipStreamSynchronize(ctx->stream));
        st->fx_valid = ctx->adj_fx;
        return SWM_OK;
    }

    // the cotangent of Fx (nux, nuy) accumulated by the last pullback
    static int download_fx(swm_ctx *ctx, void *out) {
        AdjState *st = ctx->adj;
        if (!st || !st->fx_valid) return ctx->fail(SWM_E_STATE, "no pullback with swm_adj_track_Fx on");
        Field f;
        f.origin = st->F0.lfx;
        f.base = (char *)ctx->arena + f.origin;
        f.es = ctx->es;
        f.p0 = (char *)f.base + ((size_t)OX + (size_t)OY * ctx->P) * ctx->es;
        int rc = get2d(ctx, f, out, ctx->nux, ctx->nuy, 1 - ctx->ep, 1);
        if (rc) return rc;
        HIPCHK(hipStreamSynchronize(ctx->stream));
        return SWM_OK;
    }

    // measurement: which 0 forward step, 1 taped step, 2 reverse step (kernels of one step,
    // tape in place), 3 stage adjoint (middle stage), 4 dissipation adjoint
    static int time(swm_ctx *ctx, int which, int reps, double *ms) {
        int rc = init(ctx);
        if (rc) return rc;
        AdjState *st = ctx->adj;
        if (st->snap.empty() && (rc = snap_alloc(ctx, 2))) return rc;
        st->i0 = 1; st->t0 = 0;
        if ((rc = copy_state(ctx, 0, true))) return rc;
        st->F = st->F0;
        adj::AdjFields Ft = st->F0;
        if ((rc = tape_step(ctx, 0, Ft))) return rc;
        K kl{ctx, st};
        const adj::AdjConsts<R> k = consts(ctx, 0);
        adj::StageAdjArgs<R> S{};
        S.s = Ft.s[1]; S.H = Ft.H; S.fq = Ft.fq; S.fq_2d = Ft.fq_2d; S.qend = Ft.qend; S.lqend = Ft.lqend;
        S.fu = Ft.fu; S.fv = Ft.fv; S.fu_2d = Ft.fu_2d; S.fv_2d = Ft.fv_2d;
        S.l0 = Ft.Z; S.mu_in = Ft.Ma; S.mu_x[0] = S.mu_x[1] = Ft.A; S.mu_out = Ft.Mb; S.a_out = Ft.A;
        S.a = k.a[1]; S.b = k.b[1]; S.g = k.g; S.gamma = k.gamma; S.k12 = k.k12; S.oma = k.oma; S.cont = k.cont;
        S.nx = k.nx; S.ny = k.ny; S.P = k.P;
        adj::DissAdjArgs<R> D{};
        D.s = Ft.pre; D.H = Ft.H; D.lin = Ft.L; D.out = Ft.Z;
        D.cD = k.cD; D.rD = k.rD; D.cS = k.cS; D.nuB = k.nuB; D.dtd = k.dtd; D.oma = k.oma;
        D.nx = k.nx; D.ny = k.ny; D.P = k.P;
        auto launch = [&]() -> int {
            switch (which) {
            case 0: { bool nan = false; return LF::step(ctx, 1, 0, &nan); }
            case 1: { adj::AdjFields G = st->F0; return tape_step(ctx, 0, G); }
            case 2: { st->F = Ft; return adj::reverse_step<R>(kl, st->F, k, 1); }
            case 3: return kl.stage(S, k.per, 1, 0, 0, k.lin ? adj::ADJ_LIN : k.sad ? adj::ADJ_SAD : adj::ADJ_AH);
            case 4: return kl.diss(D, k.per, k.drag, k.smag);
            }
            return ctx->fail(SWM_E_ARG, "unknown adjoint kernel %d", which);
        };
        if ((rc = launch())) return rc;
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, ctx->stream));
        for (int r = 0; r < reps && rc == SWM_OK; ++r) rc = launch();
        HIPCHK(hipEventRecord(e1, ctx->stream));
        HIPCHK(hipEventSynchronize(e1));
        float tm = 0.0f;
        HIPCHK(hipEventElapsedTime(&tm, e0, e1));
        HIPCHK(hipEventDestroy(e0));
        HIPCHK(hipEventDestroy(e1));
        *ms = (double)tm / reps;
        // the state the measurement started from
        if (!rc) rc = copy_state(ctx, 0, false);
        HIPCHK(hipStreamSynchronize(ctx->stream));
        st->recorded = false;
        return rc;
    }
};

template <class F> struct Ens {
    using Q = typename F::Q;
    using ST = typename F::ST;

    // Record step i of every member (time (i-1)*dtint of the member) into a packed plan.
    // Planning changes no member: a member whose kept q and KEu+KEv are still to be formed
    // from its state (advcor_pending) gets that update as the plan's first item and is listed
    // in pl.init; the flag is cleared when the plan is launched (launched()).
    static int plan(swm_ens *E, int64_t i, swm_ens::Plan &pl) {
        const size_t nm = E->m.size();
        std::vector<Recorder> recs(nm);
        pl = swm_ens::Plan();
        for (size_t k = 0; k < nm; ++k) {
            swm_ctx *ctx = E->m[k];
            const bool pending = ctx->advcor_pending;
            ctx->rec = &recs[k];
            Slots u0, u1;
            int rc = Launch<F>::stages(ctx, 1 << 20, (i - 1) * ctx->c.dtint, &u0, &u1);
            if (!rc && (adv_form(ctx->c) & ADV_KEEP)) {          // time_integration.jl:208-211; the quadratic
                const bool due = i % ctx->c.nstep_advcor == 0;   // drag's KEu, KEv every step (nstep_diff = 1)
                if (due || ctx->c.bottom_drag == SWM_DRAG_QUADRATIC) rc = Launch<F>::advcor(ctx, u0, due, false);
            }
            const int out = pick({ctx->pu, u0.uv});
            if (!rc) rc = Launch<F>::dissipation(ctx, u0, out);
            ctx->rec = nullptr;
            ctx->advcor_pending = pending;                      // stages() cleared it while recording
            if (rc) return E->fail(rc, "member %zu: %s", k, ctx->err.c_str());
            if (pending) pl.init.push_back((int)k);
            pl.next_pu.push_back(out);
            pl.next_pe.push_back(u0.e);
            pl.next_uv.push_back(out);
        }
        const size_t nk = recs[0].items.size();
        for (size_t k = 1; k < nm; ++k) {
            bool same = recs[k].items.size() == nk;
            for (size_t j = 0; same && j < nk; ++j)
                same = recs[k].items[j].kind == recs[0].items[j].kind && recs[k].items[j].ma == recs[0].items[j].ma &&
                       recs[k].items[j].mb == recs[0].items[j].mb;
            if (!same) return E->fail(SWM_E_ARG, "member %zu runs another kernel sequence than member 0", k);
        }
        auto align = [&]() { pl.blob.resize((pl.blob.size() + 255) & ~(size_t)255, 0); };
        for (size_t j = 0; j < nk; ++j) {
            const RecItem &it0 = recs[0].items[j];
            pl.kind.push_back(it0.kind);
            pl.ma.push_back(it0.ma);
            pl.mb.push_back(it0.mb);
            align();
            pl.args_off.push_back(pl.blob.size());
            for (size_t k = 0; k < nm; ++k) pl.blob.insert(pl.blob.end(), recs[k].items[j].args.begin(), recs[k].items[j].args.end());
            // tile table: waves per member, and (periodic stages) blocks of the q column kernel
            std::vector<int> start(1, 0), qstart(1, 0);
            for (size_t k = 0; k < nm; ++k) {
                int tiles, ny;
                if (it0.kind == 2) {
                    const DissArgs<F> *A = (const DissArgs<F> *)recs[k].items[j].args.data();
                    tiles = A->nstrips * A->nseg;
                    ny = A->ny;
                } else {
                    const StageArgs<F> *A = (const StageArgs<F> *)recs[k].items[j].args.data();
                    tiles = A->nstrips * A->nseg;
                    ny = A->ny;
                }
                start.push_back(start.back() + tiles);
                qstart.push_back(qstart.back() + (ny + 1 + 255) / 256);
            }
            align();
            pl.tab_off.push_back(pl.blob.size());
            pl.blob.insert(pl.blob.end(), (const char *)start.data(), (const char *)(start.data() + start.size()));
            align();
            pl.qtab_off.push_back(pl.blob.size());
            pl.blob.insert(pl.blob.end(), (const char *)qstart.data(), (const char *)(qstart.data() + qstart.size()));
            pl.tiles.push_back(start.back());
            pl.qblocks.push_back(qstart.back());
        }
        // the kernel sequence itself closes the plan: two steps whose arguments are the same
        // bytes but whose kernels differ (the kept fields' update with or without q) must not
        // compare equal, or a buffer or graph of the one would run for the other
        align();
        for (size_t j = 0; j < nk; ++j) {
            const int t[3] = {pl.kind[j], pl.ma[j], pl.mb[j]};
            pl.blob.insert(pl.blob.end(), (const char *)t, (const char *)(t + 3));
        }
        return SWM_OK;
    }

    // the members' kept fields formed by a launched plan are no longer pending
    static void launched(swm_ens *E, const swm_ens::Plan &pl) {
        for (int k : pl.init) E->m[k]->advcor_pending = false;
    }

    // Make buffer b hold plan pl (upload only when it differs).  The captured graph holds the
    // launch sizes of both buffers' plans and the per-member items' arguments by value, so an
    // upload invalidates it.
    static int load(swm_ens *E, int b, const swm_ens::Plan &pl) {
        swm_ens::Buf &B = E->buf[b];
        if (B.valid && B.plan.blob == pl.blob) return SWM_OK;
        E->graph_ok = false;
        if (hipStreamSynchronize(E->stream) != hipSuccess) return E->fail(SWM_E_HIP, "stream synchronize");
        if (pl.blob.size() > B.cap) {
            if (B.dev) (void)hipFree(B.dev);
            B.cap = std::max<size_t>(pl.blob.size() * 2, 1 << 16);
            if (hipMalloc(&B.dev, B.cap) != hipSuccess) return E->fail(SWM_E_HIP, "ensemble plan buffer");
        }
        if (hipMemcpy(B.dev, pl.blob.data(), pl.blob.size(), hipMemcpyHostToDevice) != hipSuccess)
            return E->fail(SWM_E_HIP, "ensemble plan upload");
        B.plan = pl;
        B.valid = true;
        return SWM_OK;
    }

    // the batched launches of buffer b's step on the ensemble stream
    template <bool PER> static int launch(swm_ens *E, int b) {
        const swm_ens::Buf &B = E->buf[b];
        const swm_ens::Plan &pl = B.plan;
        const int nm = (int)E->m.size();
        const swm_config &c0 = E->m[0]->c;
        char *dev = (char *)B.dev;
        const dim3 block(WAVE * WAVES_PER_BLOCK);
        // a periodic pair's seam strips (kind 5, recorded right after the pair's other strips,
        // kind 4) are issued first, on the second stream, so that their few long marches start
        // before the other strips fill the device
        const size_t nk = pl.kind.size();
        std::vector<size_t> order;
        order.reserve(nk);
        for (size_t j = 0; j < nk; ++j) {
            if (pl.kind[j] == 4 && j + 1 < nk && pl.kind[j + 1] == 5) {
                order.push_back(j + 1);
                order.push_back(j);
                ++j;
            } else
                order.push_back(j);
        }
        for (size_t q = 0; q < nk; ++q) {
            const size_t j = order[q];
            const int *tab = (const int *)(dev + pl.tab_off[j]), *qtab = (const int *)(dev + pl.qtab_off[j]);
            const dim3 grid(nblocks_waves(pl.tiles[j]));
            if (pl.kind[j] == 0) {
                const StageArgs<F> *a = (const StageArgs<F> *)(dev + pl.args_off[j]);
                const int adv = adv_form(c0);                     // one form for all members
                if (PER && adv == ADV_AH) {
                    if (pl.ma[j] == M_RK_FIRST) hipLaunchKernelGGL((qend_batch<F, Q>), dim3(pl.qblocks[j]), dim3(256), 0, E->stream, a, qtab, nm);
                    else hipLaunchKernelGGL((qend_batch<F, ST>), dim3(pl.qblocks[j]), dim3(256), 0, E->stream, a, qtab, nm);
                }
#define SWM_SB(M_)                                                                                                  \
    if (adv == ADV_SAD) hipLaunchKernelGGL((stage_batch<F, PER, M_, ADV_SAD>), grid, block, 0, E->stream, a, tab, nm);  \
    else if (adv == ADV_LIN) hipLaunchKernelGGL((stage_batch<F, PER, M_, ADV_LIN>), grid, block, 0, E->stream, a, tab, nm); \
    else if (adv == (ADV_AH | ADV_KEEP)) hipLaunchKernelGGL((stage_batch<F, PER, M_, ADV_AH | ADV_KEEP>), grid, block, 0, E->stream, a, tab, nm); \
    else if (adv == (ADV_SAD | ADV_KEEP)) hipLaunchKernelGGL((stage_batch<F, PER, M_, ADV_SAD | ADV_KEEP>), grid, block, 0, E->stream, a, tab, nm); \
    else hipLaunchKernelGGL((stage_batch<F, PER, M_>), grid, block, 0, E->stream, a, tab, nm);
                switch (pl.ma[j]) {
                case M_RK_FIRST: SWM_SB(M_RK_FIRST) break;
                case M_RK_MID: SWM_SB(M_RK_MID) break;
                case M_RK_LAST: SWM_SB(M_RK_LAST) break;
                default: return E->fail(SWM_E_UNSUPPORTED, "stage mode %d in an ensemble", pl.ma[j]);
                }
#undef SWM_SB
            } else if (pl.kind[j] == 1 && adv_form(c0) != ADV_AH) {   // Sadourny / linear members (RK4, T = Tprog)
                const StageArgs<F> *a = (const StageArgs<F> *)(dev + pl.args_off[j]);
                if constexpr (!F::mixed) {
                    const int adv = adv_form(c0);
                    if (adv != ADV_SAD && adv != ADV_LIN && adv != (ADV_AH | ADV_KEEP) && adv != (ADV_SAD | ADV_KEEP))
                        return E->fail(SWM_E_UNSUPPORTED, "right-hand side form %d of fused stages in an ensemble", adv);
#define SWM_FBV(MA_, MB_)                                                                                          \
    if (adv == ADV_SAD) hipLaunchKernelGGL((fused_batch<F, MA_, MB_, false, PER, ADV_SAD>), grid, block, 0, E->stream, a, tab, nm); \
    else if (adv == ADV_LIN) hipLaunchKernelGGL((fused_batch<F, MA_, MB_, false, PER, ADV_LIN>), grid, block, 0, E->stream, a, tab, nm); \
    else if (adv == (ADV_AH | ADV_KEEP)) hipLaunchKernelGGL((fused_batch<F, MA_, MB_, false, PER, ADV_AH | ADV_KEEP>), grid, block, 0, E->stream, a, tab, nm); \
    else hipLaunchKernelGGL((fused_batch<F, MA_, MB_, false, PER, ADV_SAD | ADV_KEEP>), grid, block, 0, E->stream, a, tab, nm)
                    if (pl.ma[j] == M_RK_FIRST && pl.mb[j] == M_RK_MID) { SWM_FBV(M_RK_FIRST, M_RK_MID); }
                    else if (pl.ma[j] == M_RK_MID && pl.mb[j] == M_RK_LAST) { SWM_FBV(M_RK_MID, M_RK_LAST); }
                    else return E->fail(SWM_E_UNSUPPORTED, "fused stages %d+%d in an ensemble", pl.ma[j], pl.mb[j]);
#undef SWM_FBV
                } else {
                    return E->fail(SWM_E_UNSUPPORTED, "fused stages of this right-hand side with mixed formats");
                }
            } else if (pl.kind[j] == 1) {
                const StageArgs<F> *a = (const StageArgs<F> *)(dev + pl.args_off[j]);
                const bool cs = c0.continuity != SWM_CONT_ITSELF;
#define SWM_FB(MA_, MB_)                                                                                        \
    if (cs) hipLaunchKernelGGL((fused_batch<F, MA_, MB_, true, PER>), grid, block, 0, E->stream, a, tab, nm);   \
    else hipLaunchKernelGGL((fused_batch<F, MA_, MB_, false, PER>), grid, block, 0, E->stream, a, tab, nm)
                if (pl.ma[j] == M_RK_FIRST && pl.mb[j] == M_RK_MID) SWM_FB(M_RK_FIRST, M_RK_MID);
                else if (pl.ma[j] == M_RK_MID && pl.mb[j] == M_RK_LAST) SWM_FB(M_RK_MID, M_RK_LAST);
                else if (pl.ma[j] == M_RK_FIRST && pl.mb[j] == M_RK_LAST) SWM_FB(M_RK_FIRST, M_RK_LAST);
                else return E->fail(SWM_E_UNSUPPORTED, "fused stages %d+%d in an ensemble", pl.ma[j], pl.mb[j]);
#undef SWM_FB
            } else if (pl.kind[j] == 4 || pl.kind[j] == 5) {      // periodic pairs: strips clear of the seam / at it
                const StageArgs<F> *a = (const StageArgs<F> *)(dev + pl.args_off[j]);
                const bool cs = c0.continuity != SWM_CONT_ITSELF;
                // the seam strips read what the other strips read and write other cells: they run
                // on the second stream beside them, forked before the seam strips and joined after
                // the other strips (graph capture keeps both branches)
                hipStream_t st = E->stream;
                const bool joins = pl.kind[j] == 4 ? j + 1 < nk && pl.kind[j + 1] == 5 : !(j > 0 && pl.kind[j - 1] == 4);
                if (pl.kind[j] == 5) {
                    if (hipEventRecord(E->fork, E->stream) != hipSuccess || hipStreamWaitEvent(E->stream2, E->fork, 0) != hipSuccess)
                        return E->fail(SWM_E_HIP, "fork");
                    st = E->stream2;
                }
                if (pl.tiles[j] > 0) {
#define SWM_FB2(MA_, MB_)                                                                                   \
    if (pl.kind[j] == 4) {                                                                                  \
        if (cs) hipLaunchKernelGGL((fused_batch<F, MA_, MB_, true, true>), grid, block, 0, st, a, tab, nm); \
        else hipLaunchKernelGGL((fused_batch<F, MA_, MB_, false, true>), grid, block, 0, st, a, tab, nm);  \
    } else if (cs) hipLaunchKernelGGL((fused_seam_batch<F, MA_, MB_, true>), grid, block, 0, st, a, tab, nm); \
    else hipLaunchKernelGGL((fused_seam_batch<F, MA_, MB_, false>), grid, block, 0, st, a, tab, nm)
                    if (pl.ma[j] == M_RK_FIRST && pl.mb[j] == M_RK_MID) { SWM_FB2(M_RK_FIRST, M_RK_MID); }
                    else if (pl.ma[j] == M_RK_MID && pl.mb[j] == M_RK_LAST) { SWM_FB2(M_RK_MID, M_RK_LAST); }
                    else if (pl.ma[j] == M_RK_FIRST && pl.mb[j] == M_RK_LAST) { SWM_FB2(M_RK_FIRST, M_RK_LAST); }
                    else return E->fail(SWM_E_UNSUPPORTED, "fused stages %d+%d in an ensemble", pl.ma[j], pl.mb[j]);
#undef SWM_FB2
                }
                if (joins && (hipEventRecord(E->join, E->stream2) != hipSuccess ||
                              hipStreamWaitEvent(E->stream, E->join, 0) != hipSuccess))
                    return E->fail(SWM_E_HIP, "join");
            } else if (pl.kind[j] == 2) {
                const DissArgs<F> *a = (const DissArgs<F> *)(dev + pl.args_off[j]);
                const bool smag = c0.diffusion == SWM_DIFF_SMAGORINSKY;
#define SWM_DB(D_, S_) hipLaunchKernelGGL((dissipation_batch<F, PER, D_, S_>), grid, block, 0, E->stream, a, tab, nm)
                if (c0.bottom_drag == SWM_DRAG_QUADRATIC) { if (smag) SWM_DB(0, true); else SWM_DB(0, false); }
                else if (c0.bottom_drag == SWM_DRAG_LINEAR) { if (smag) SWM_DB(1, true); else SWM_DB(1, false); }
                else { if (smag) SWM_DB(2, true); else SWM_DB(2, false); }
#undef SWM_DB
            } else if (pl.kind[j] == 6) {                        // kept q and KEu+KEv: one launch per member
                for (int k = 0; k < nm; ++k) {
                    StageArgs<F> A;
                    memcpy(&A, pl.blob.data() + pl.args_off[j] + (size_t)k * sizeof A, sizeof A);
                    const swm_config &c = E->m[k]->c;
                    const dim3 g((c.nx + 2 + 255) / 256, c.ny + 2), b(256);
                    if (pl.ma[j] & 2) hipLaunchKernelGGL((advcor_kernel<F, Q, true, true>), g, b, 0, E->stream, A);
                    else if (pl.ma[j] & 1) hipLaunchKernelGGL((advcor_kernel<F, Q, true, false>), g, b, 0, E->stream, A);
                    else hipLaunchKernelGGL((advcor_kernel<F, Q, false, false>), g, b, 0, E->stream, A);
                }
            } else return E->fail(SWM_E_UNSUPPORTED, "kernel kind %d in an ensemble", pl.kind[j]);
        }
        if (hipGetLastError() != hipSuccess) return E->fail(SWM_E_HIP, "batched launch");
        return SWM_OK;
    }
    static int launch_buf(swm_ens *E, int b) { return E->per ? launch<true>(E, b) : launch<false>(E, b); }

    // host work a member has at step i: tracer snapshot or advection, NaN check
    static bool events(const swm_ens *E, int64_t i) {
        for (const swm_ctx *ctx : E->m) {
            const swm_config &c = ctx->c;
            if (c.tracer_advection && ((i + c.nadvstep_half) % c.nadvstep == 0 || i % c.nadvstep == 0)) return true;
            if (c.nout > 0 && i % c.nout == 0) return true;
        }
        return false;
    }

    static int capture(swm_ens *E) {
        if (E->exec) { (void)hipGraphExecDestroy(E->exec); E->exec = nullptr; }
        if (E->graph) { (void)hipGraphDestroy(E->graph); E->graph = nullptr; }
        if (hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return E->fail(SWM_E_HIP, "graph capture");
        int rc = launch_buf(E, 0);
        if (!rc) rc = launch_buf(E, 1);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(E->stream, &g);
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) return E->fail(SWM_E_HIP, "graph capture end");
        E->graph = g;
        if (hipGraphInstantiate(&E->exec, g, nullptr, nullptr, 0) != hipSuccess) return E->fail(SWM_E_HIP, "graph instantiate");
        E->graph_ok = true;
        return SWM_OK;
    }

    static void advance(swm_ens *E, const swm_ens::Plan &pl) {
        for (size_t k = 0; k < E->m.size(); ++k) {
            E->m[k]->pu = pl.next_pu[k];
            E->m[k]->pe = pl.next_pe[k];
        }
    }

    // ------------------------------------------------------------- persistent launch
    static constexpr const char *WHY_SETTLE = "the plans do not repeat every two steps";
    static constexpr int PHASE_TRACE = 8192;              // barrier exit times recorded when timing phases
    //
    // Steps whose plans repeat every two steps and whose members all run the fused RK4 pairs
    // with T = Tprog, the continuity itself and Arakawa-Hsu advection run in one cooperative
    // launch (ens_persist_kernel): the plans of
    // the two parities in the plan buffers, each member's tracer and NaN-check schedule in a
    // record.  *ran is false (and nothing changed) when the form does not allow it.
    static int persist(swm_ens *E, int64_t nsteps, int64_t i0, int64_t *done, int64_t *nan_steps, bool *ran) {
        *ran = false;
        E->persist_why = "";
        auto skip = [&](const char *why) { E->persist_why = why; return SWM_OK; };
        if constexpr (F::mixed) {
            return skip("mixed number formats (no fused stage pairs)");
        } else {
            const size_t nm = E->m.size();
            if (!E->persistent) return skip("turned off");
            if (nsteps < 1) return SWM_OK;
            for (const swm_ctx *ctx : E->m) {
                const swm_config &c = ctx->c;
                if (ctx->advcor_pending || adv_form(c) != ADV_AH || c.continuity != SWM_CONT_ITSELF || c.nstep_diff != 1)
                    return skip("right-hand side or continuity form");
            }
            // the plans of steps i0, i0+1 and i0+2: the third must repeat the first
            swm_ens::Plan pl[3];
            std::vector<int> pu(nm), pe(nm);
            for (size_t q = 0; q < nm; ++q) { pu[q] = E->m[q]->pu; pe[q] = E->m[q]->pe; }
            int rc = SWM_OK;
            for (int k = 0; k < 3 && !rc; ++k) {
                rc = plan(E, i0 + k, pl[k]);
                if (!rc) advance(E, pl[k]);
            }
            for (size_t q = 0; q < nm; ++q) { E->m[q]->pu = pu[q]; E->m[q]->pe = pe[q]; }
            if (rc) return rc;
            if (pl[2].blob != pl[0].blob) return skip(WHY_SETTLE);
            for (int k = 0; k < 2; ++k) {
                const swm_ens::Plan &P = pl[k];
                if (!P.init.empty() || P.kind.empty() || P.kind.size() > (size_t)PERSIST_ITEMS)
                    return skip("pending kept-field update, or too many kernels per step");
                for (size_t j = 0; j < P.kind.size(); ++j) {
                    const int kd = P.kind[j];
                    const bool pair = (P.ma[j] == M_RK_FIRST && P.mb[j] == M_RK_MID) || (P.ma[j] == M_RK_MID && P.mb[j] == M_RK_LAST);
                    if (!(kd == 2 || ((kd == 1 || kd == 4 || (kd == 5 && E->per)) && pair)))
                        return skip("kernels other than fused RK4 pairs and the dissipation");
                }
            }
            if ((rc = load(E, 0, pl[0])) || (rc = load(E, 1, pl[1]))) return rc;
            E->last = 1;
            // kernel arguments: items, barriers between phases (the strips clear of a seam and
            // the seam strips of one pair write disjoint cells and share a phase)
            PArgs<F> A;
            std::memset((void *)&A, 0, sizeof A);
            PItem items[2][PERSIST_ITEMS];
            std::memset((void *)items, 0, sizeof items);
            for (int k = 0; k < 2; ++k) {
                const swm_ens::Plan &P = pl[k];
                A.plan[k] = (const char *)E->buf[k].dev;
                A.nitems[k] = (int)P.kind.size();
                for (size_t j = 0; j < P.kind.size(); ++j) {
                    PItem &it = items[k][j];
                    it.kind = P.kind[j]; it.ma = P.ma[j]; it.mb = P.mb[j];
                    it.args = (uint32_t)P.args_off[j]; it.tab = (uint32_t)P.tab_off[j]; it.tiles = P.tiles[j];
                    const bool joined = P.kind[j] == 4 && j + 1 < P.kind.size() && P.kind[j + 1] == 5 &&
                                        P.ma[j + 1] == P.ma[j] && P.mb[j + 1] == P.mb[j];
                    it.sync = joined ? 0 : 1;
                }
            }
            // member records
            std::vector<PMember<F>> mem(nm);
            const int ntrace = E->persistent == 2 ? PHASE_TRACE : 0;
            const size_t nan_off = (sizeof(PCtl) + 255) & ~(size_t)255;
            const size_t mem_off = nan_off + ((nm * sizeof(int) + 255) & ~(size_t)255);
            const size_t trace_off = mem_off + ((nm * sizeof(PMember<F>) + 255) & ~(size_t)255);
            const size_t item_off = trace_off + (((size_t)ntrace * sizeof(long long) + 255) & ~(size_t)255);
            const size_t need = item_off + 2 * PERSIST_ITEMS * sizeof(PItem);
            if (need > E->pcap) {
                if (E->pdev) (void)hipFree(E->pdev);
                E->pdev = nullptr;
                E->pcap = std::max<size_t>(need, 1 << 16);
                if (hipMalloc(&E->pdev, E->pcap) != hipSuccess) { E->pcap = 0; return E->fail(SWM_E_HIP, "persistent launch records"); }
            }
            char *dv = (char *)E->pdev;
            PCtl *dctl = (PCtl *)dv;
            int *dnan = (int *)(dv + nan_off);
            PMember<F> *dmem = (PMember<F> *)(dv + mem_off);
            long long *dtrace = (long long *)(dv + trace_off);
            PItem *ditem = (PItem *)(dv + item_off);
            for (size_t q = 0; q < nm; ++q) {
                swm_ctx *ctx = E->m[q];
                const swm_config &c = ctx->c;
                PMember<F> &M = mem[q];
                std::memset((void *)&M, 0, sizeof M);
                M.tr = Launch<F>::tracer_args(ctx, pl[0].next_uv[q], 0, 1);
                for (int k = 0; k < 2; ++k) {
                    const int uv = pl[k].next_uv[q];
                    M.tu[k] = ctx->at<Q>(ctx->u[uv]);
                    M.tv[k] = ctx->at<Q>(ctx->v[uv]);
                    M.su[k] = (const Q *)ctx->u[uv].base;
                    M.sv[k] = (const Q *)ctx->v[uv].base;
                    const int su = k == 0 ? pu[q] : pl[0].next_pu[q], se = k == 0 ? pe[q] : pl[0].next_pe[q];
                    NanArgs &N = M.nan[k];
                    N.f[0] = ctx->at<Q>(ctx->u[su]); N.x0[0] = -1 - ctx->ep; N.x1[0] = ctx->nux + 2 - ctx->ep; N.y0[0] = -1; N.y1[0] = c.ny + 2;
                    N.f[1] = ctx->at<Q>(ctx->v[su]); N.x0[1] = -1; N.x1[1] = c.nx + 2; N.y0[1] = -1; N.y1[1] = c.ny + 1;
                    N.f[2] = ctx->at<Q>(ctx->e[se]); N.x0[2] = 0; N.x1[2] = c.nx + 1; N.y0[2] = 0; N.y1[2] = c.ny + 1;
                    N.f[3] = nullptr; N.x0[3] = 0; N.x1[3] = c.nx + 1; N.y0[3] = 1; N.y1[3] = c.ny;
                    N.P = ctx->P;
                    N.nf = 4;
                }
                M.sst[0] = ctx->at<Q>(ctx->sst[0]);
                M.sst[1] = ctx->at<Q>(ctx->sst[1]);
                M.um = (ST *)ctx->um.base;
                M.vm = (ST *)ctx->vm.base;
                M.elems = (long long)ctx->elems;
                M.tracer = c.tracer_advection ? 1 : 0;
                M.nadvstep = c.tracer_advection ? (int)c.nadvstep : 1;
                M.nout = c.nout > 0 ? (int)c.nout : 0;
                M.r_adv = (int)(i0 % M.nadvstep);
                M.r_half = (int)((i0 + (c.tracer_advection ? c.nadvstep_half : 0)) % M.nadvstep);
                M.r_nan = M.nout > 0 ? (int)(i0 % M.nout) : 0;
                M.rsst0 = ctx->rsst;
                M.nanflag = dnan + q;
            }
            if (hipMemcpyAsync(dmem, mem.data(), nm * sizeof(PMember<F>), hipMemcpyHostToDevice, E->stream) != hipSuccess)
                return E->fail(SWM_E_HIP, "persistent launch records");
            A.item = ditem;
            A.mem = dmem;
            A.nm = (int)nm;
            A.ctl = dctl;
            A.ntrace = ntrace;
            A.trace = dtrace;
            {
                const swm_config &c0 = E->m[0]->c;      // members share drag and diffusion
                const int drag = c0.bottom_drag == SWM_DRAG_QUADRATIC ? 0 : c0.bottom_drag == SWM_DRAG_LINEAR ? 1 : 2;
                A.diss = 2 * drag + (c0.diffusion == SWM_DIFF_SMAGORINSKY ? 0 : 1);
            }
            // the launch: as many blocks as stay resident, a multiple of the XCD count
            auto kern = E->per ? (const void *)&ens_persist_kernel<F, true> : (const void *)&ens_persist_kernel<F, false>;
            int dev = 0, ncu = 0, per_cu = 0, coop = 0, clk_khz = 0;
            if (hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
                hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess ||
                hipDeviceGetAttribute(&clk_khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess ||
                hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, WAVE * WAVES_PER_BLOCK, 0) != hipSuccess)
                return E->fail(SWM_E_HIP, "persistent launch: device attributes");
            const int nb = (per_cu * ncu) / NXCD * NXCD;
            if (!coop || nb < NXCD) return skip("no cooperative launch on this device");
            A.timeout = (long long)clk_khz * 1000 * 5;       // 5 s of the wall clock
            int64_t left = nsteps, at = i0, stopped = -1;
            std::vector<int> nanf(nm, 0);
            while (left > 0 && stopped < 0) {
                const int chunk = (int)std::min<int64_t>(left, 1 << 20);
                A.nsteps = chunk;
                // members' residues at the chunk's first step
                if (at != i0) {
                    for (size_t q = 0; q < nm; ++q) {
                        const swm_config &c = E->m[q]->c;
                        mem[q].r_adv = (int)(at % mem[q].nadvstep);
                        mem[q].r_half = (int)((at + (c.tracer_advection ? c.nadvstep_half : 0)) % mem[q].nadvstep);
                        mem[q].r_nan = mem[q].nout > 0 ? (int)(at % mem[q].nout) : 0;
                        mem[q].rsst0 = E->m[q]->rsst;
                    }
                    if (hipMemcpyAsync(dmem, mem.data(), nm * sizeof(PMember<F>), hipMemcpyHostToDevice, E->stream) != hipSuccess)
                        return E->fail(SWM_E_HIP, "persistent launch records");
                }
                PCtl c0{};
                c0.stop = 0x7fffffff;
                if (hipMemcpyAsync(dctl, &c0, sizeof c0, hipMemcpyHostToDevice, E->stream) != hipSuccess ||
                    hipMemcpyAsync(ditem, items, sizeof items, hipMemcpyHostToDevice, E->stream) != hipSuccess ||
                    hipMemsetAsync(dnan, 0, nm * sizeof(int), E->stream) != hipSuccess)
                    return E->fail(SWM_E_HIP, "persistent launch control");
                void *args[] = {&A};
                if (hipLaunchCooperativeKernel(kern, dim3(nb), dim3(WAVE * WAVES_PER_BLOCK), args, 0, E->stream) != hipSuccess)
                    return E->fail(SWM_E_HIP, "persistent launch");
                PCtl c1{};
                if (hipMemcpyAsync(&c1, dctl, sizeof c1, hipMemcpyDeviceToHost, E->stream) != hipSuccess ||
                    hipMemcpyAsync(nanf.data(), dnan, nm * sizeof(int), hipMemcpyDeviceToHost, E->stream) != hipSuccess ||
                    hipStreamSynchronize(E->stream) != hipSuccess)
                    return E->fail(SWM_E_HIP, "persistent launch");
                if (c1.abort) return E->fail(SWM_E_HIP, "persistent launch: a grid barrier timed out (members' states undefined)");
                const int ran_steps = c1.stop < chunk ? c1.stop + 1 : chunk;
                if (ntrace > 0 && at == i0) {
                    std::vector<long long> tr(ntrace);
                    if (hipMemcpy(tr.data(), dtrace, ntrace * sizeof(long long), hipMemcpyDeviceToHost) != hipSuccess)
                        return E->fail(SWM_E_HIP, "persistent launch trace");
                    phase_times(E, A, items, mem, tr, ran_steps, clk_khz);
                }
                // the members' slots and tracer slot after the steps that ran
                for (size_t q = 0; q < nm; ++q) {
                    swm_ctx *ctx = E->m[q];
                    const PMember<F> &M = mem[q];
                    const bool nan = c1.stop < chunk && nanf[q];
                    // steps since i0 whose copy-back happened: the slots repeat every two steps
                    const int64_t acc = (at - i0) + (nan ? ran_steps - 1 : ran_steps);
                    ctx->pu = (acc & 1) ? pl[0].next_pu[q] : pu[q];
                    ctx->pe = (acc & 1) ? pl[0].next_pe[q] : pe[q];
                    if (M.tracer) {
                        const int a0 = (M.nadvstep - M.r_adv) % M.nadvstep;
                        const int n = ran_steps > a0 ? (ran_steps - 1 - a0) / M.nadvstep + 1 : 0;
                        ctx->rsst = M.rsst0 ^ (n & 1);
                    }
                    if (nan) nan_steps[q] = at + c1.stop;
                }
                if (c1.stop < chunk) stopped = at + c1.stop;
                at += ran_steps;
                left -= ran_steps;
                *done += ran_steps;
                // the next chunk starts at the parity reached
                if (left > 0 && stopped < 0 && ((at - i0) & 1)) {
                    std::swap(A.plan[0], A.plan[1]);
                    std::swap(items[0], items[1]);
                    std::swap(A.nitems[0], A.nitems[1]);
                    for (PMember<F> &M : mem) {
                        std::swap(M.tu[0], M.tu[1]); std::swap(M.tv[0], M.tv[1]);
                        std::swap(M.su[0], M.su[1]); std::swap(M.sv[0], M.sv[1]);
                        std::swap(M.nan[0], M.nan[1]);
                    }
                }
            }
            E->persist_steps += *done;
            *ran = true;
            return SWM_OK;
        }
    }

    // Average duration of the phases of the persistent launch's steps from block 0's barrier
    // exit times: the first fused pair, the second, the dissipation (each with its barrier),
    // over the steps that follow a step without tracer or NaN-check work; and whole steps.
    static void phase_times(swm_ens *E, const PArgs<F> &A, const PItem (&items)[2][PERSIST_ITEMS], const std::vector<PMember<F>> &mem,
                            const std::vector<long long> &tr, int steps, int clk_khz) {
        double sum[4] = {0, 0, 0, 0};
        int cnt[4] = {0, 0, 0, 0};
        int k = 0;                                        // barriers so far
        bool clean = false;                               // the previous step had no events
        const int ntr = (int)tr.size();
        auto ms = [&](int b) { return (double)(tr[b] - tr[b - 1]) / clk_khz; };
        for (int s = 0; s < steps && k + 8 < ntr; ++s) {
            const int p = s & 1, k0 = k;
            for (int j = 0; j < A.nitems[p]; ++j) {
                const PItem &it = items[p][j];
                if (!it.sync) continue;
                ++k;
                const int ph = it.kind == 2 ? 2 : it.ma == M_RK_FIRST ? 0 : 1;
                if (clean) { sum[ph] += ms(k); cnt[ph] += 1; }
            }
            bool snap = false, adv = false, both = false, nan = false;
            for (const PMember<F> &M : mem) {
                const bool sn = M.tracer && (M.r_half + s) % M.nadvstep == 0, ad = M.tracer && (M.r_adv + s) % M.nadvstep == 0;
                snap |= sn; adv |= ad; both |= sn && ad;
                nan |= M.nout > 0 && (M.r_nan + s) % M.nout == 0;
            }
            k += (snap && both) + (adv && nan) + nan;
            if (clean && !(snap || adv || nan) && k0 > 0) {
                sum[3] += (double)(tr[k] - tr[k0]) / clk_khz;
                cnt[3] += 1;
            }
            clean = !(snap || adv || nan);
        }
        for (int q = 0; q < 4; ++q) E->phase_ms[q] = cnt[q] ? sum[q] / cnt[q] : -1.0;
    }

    static int step(swm_ens *E, int64_t nsteps, int64_t i0, int64_t *done, int64_t *nan_steps, int64_t *gsteps) {
        const size_t nm = E->m.size();
        *done = 0;
        *gsteps = 0;
        E->persist_steps = 0;
        for (size_t k = 0; k < nm; ++k) nan_steps[k] = 0;
        bool ran = false;
        int rc = persist(E, nsteps, i0, done, nan_steps, &ran);
        if (rc) return rc;
        if (ran) return check_tracer_flags(E);
        if (E->persist_why == WHY_SETTLE && nsteps > 1) {
            // a fresh state: the slot roles settle into their two-step cycle after one step
            if ((rc = batched(E, 1, i0, done, nan_steps, gsteps))) return rc;
            if (*done < 1) return check_tracer_flags(E);
            for (size_t k = 0; k < nm; ++k)
                if (nan_steps[k]) return check_tracer_flags(E);
            int64_t d = 0;
            if ((rc = persist(E, nsteps - 1, i0 + 1, &d, nan_steps, &ran))) return rc;
            if (ran) {
                *done += d;
                return check_tracer_flags(E);
            }
            int64_t g = 0;
            rc = batched(E, nsteps - 1, i0 + 1, &d, nan_steps, &g);
            *done += d;
            *gsteps += g;
            return rc;
        }
        return batched(E, nsteps, i0, done, nan_steps, gsteps);
    }

    // steps as batched launches, pairs of steps without host work as graph replays
    static int batched(swm_ens *E, int64_t nsteps, int64_t i0, int64_t *done, int64_t *nan_steps, int64_t *gsteps) {
        const size_t nm = E->m.size();
        *done = 0;
        *gsteps = 0;
        for (size_t k = 0; k < nm; ++k) nan_steps[k] = 0;
        swm_ens::Plan p0, p1;
        for (int64_t k = 0; k < nsteps;) {
            const int64_t i = i0 + k;
            int rc = plan(E, i, p0);
            if (rc) return rc;
            // two steps as one graph replay: both plans already in the buffers, no host work
            if (k + 1 < nsteps && !events(E, i) && !events(E, i + 1) && E->buf[0].valid && E->buf[1].valid &&
                p0.blob == E->buf[0].plan.blob) {
                std::vector<int> pu(nm), pe(nm);
                for (size_t q = 0; q < nm; ++q) { pu[q] = E->m[q]->pu; pe[q] = E->m[q]->pe; }
                advance(E, p0);                                  // step i+1 planned as after step i
                launched(E, p0);
                rc = plan(E, i + 1, p1);
                for (size_t q = 0; q < nm; ++q) { E->m[q]->pu = pu[q]; E->m[q]->pe = pe[q]; }
                for (int q : p0.init) E->m[q]->advcor_pending = true;
                if (rc) return rc;
                if (p1.blob == E->buf[1].plan.blob) {
                    if (!E->graph_ok && (rc = capture(E))) return rc;
                    if (hipGraphLaunch(E->exec, E->stream) != hipSuccess) return E->fail(SWM_E_HIP, "graph launch");
                    advance(E, p0);
                    advance(E, p1);
                    launched(E, p0);
                    launched(E, p1);
                    k += 2;
                    *gsteps += 2;
                    *done = k;
                    continue;
                }
            }
            // one eager step: the plan into the buffer it matches, else the one not used last
            int b = E->buf[0].valid && E->buf[0].plan.blob == p0.blob ? 0
                  : E->buf[1].valid && E->buf[1].plan.blob == p0.blob ? 1 : 1 - E->last;
            if ((rc = load(E, b, p0))) return rc;
            E->last = b;
            if ((rc = launch_buf(E, b))) return rc;
            launched(E, p0);
            // per-member host work, in the member's own order (time_integration.jl:226-237)
            bool any_nan = false;
            for (size_t q = 0; q < nm; ++q) {
                swm_ctx *ctx = E->m[q];
                if ((rc = Launch<F>::tracer(ctx, i, p0.next_uv[q]))) return E->fail(rc, "member %zu: %s", q, ctx->err.c_str());
                if (ctx->c.nout > 0 && i % ctx->c.nout == 0) {
                    bool nan = false;
                    if ((rc = Launch<F>::nan_check(ctx, ctx->pu, ctx->pe, ctx->rsst, &nan)))
                        return E->fail(rc, "member %zu: %s", q, ctx->err.c_str());
                    if (nan) { nan_steps[q] = i; any_nan = true; }
                }
            }
            for (size_t q = 0; q < nm; ++q)
                if (!nan_steps[q]) { E->m[q]->pu = p0.next_pu[q]; E->m[q]->pe = p0.next_pe[q]; }
            ++k;
            *done = k;
            if (any_nan) break;
        }
        return check_tracer_flags(E);
    }

    // a tracer departure point that was NaN or outside the grid in any member
    static int check_tracer_flags(swm_ens *E) {
        const size_t nm = E->m.size();
        if (hipStreamSynchronize(E->stream) != hipSuccess) return E->fail(SWM_E_HIP, "stream synchronize");
        int flags[2] = {0, 0};
        for (size_t q = 0; q < nm; ++q) {
            swm_ctx *ctx = E->m[q];
            if (hipMemcpy(flags, ctx->dflags, sizeof flags, hipMemcpyDeviceToHost) != hipSuccess) return E->fail(SWM_E_HIP, "flags");
            if (flags[1]) {
                int z = 0;
                (void)hipMemcpy(ctx->dflags + 1, &z, sizeof z, hipMemcpyHostToDevice);
                return E->fail(SWM_E_ARG, "member %zu: tracer departure point is NaN or outside the grid", q);
            }
        }
        return SWM_OK;
    }

    // Time `reps` batched steps (plan of the current state, into buffer 0) with HIP events.
    // The plan leaves out pending updates of the kept fields from the initial state, so that
    // every candidate of the autotuner times the same kernel sequence.
    static int time_step(swm_ens *E, int reps, double *ms) {
        swm_ens::Plan pl;
        std::vector<char> pending;
        for (swm_ctx *ctx : E->m) {
            pending.push_back(ctx->advcor_pending);
            ctx->advcor_pending = false;
        }
        int rc = plan(E, 1, pl);
        for (size_t k = 0; k < E->m.size(); ++k) E->m[k]->advcor_pending = pending[k] != 0;
        if (rc) return rc;
        E->buf[0].valid = false;
        if ((rc = load(E, 0, pl))) return rc;
        if ((rc = launch_buf(E, 0))) return rc;                 // warm-up
        hipEvent_t e0, e1;
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return E->fail(SWM_E_HIP, "event");
        (void)hipEventRecord(e0, E->stream);
        for (int r = 0; r < reps && !rc; ++r) rc = launch_buf(E, 0);
        (void)hipEventRecord(e1, E->stream);
        (void)hipEventSynchronize(e1);
        float t = 0;
        (void)hipEventElapsedTime(&t, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *ms = t / reps;
        return rc;
    }

    // Autotuner: for each kernel class (stage, fused pair, dissipation) try segment lengths
    // and keep the one with the fastest batched step; the others held at their choice.
    // The timed steps leave each member's state in its slots, but they also update the kept
    // q and KEu+KEv (nstep_advcor > 0) and take the pending update from the initial state:
    // both are saved before timing and put back after it.
    static int tune(swm_ens *E, int reps, int *Ls, int *Lf, int *Ld) {
        struct Kept { swm_ctx *ctx; char *buf; size_t n; bool pending; };
        std::vector<Kept> kept;
        int rc = hipStreamSynchronize(E->stream) == hipSuccess ? SWM_OK : E->fail(SWM_E_HIP, "stream synchronize");
        for (swm_ctx *ctx : E->m) {
            if (rc || !(adv_form(ctx->c) & ADV_KEEP)) continue;
            Kept k{ctx, nullptr, ctx->elems * (size_t)ctx->qs.es, ctx->advcor_pending};
            if (hipMalloc((void **)&k.buf, 2 * k.n) != hipSuccess) { rc = E->fail(SWM_E_HIP, "kept fields' copy"); break; }
            kept.push_back(k);
            if (hipMemcpy(k.buf, ctx->qs.base, k.n, hipMemcpyDeviceToDevice) != hipSuccess ||
                hipMemcpy(k.buf + k.n, ctx->ks.base, k.n, hipMemcpyDeviceToDevice) != hipSuccess)
                rc = E->fail(SWM_E_HIP, "kept fields' copy");
        }
        if (!rc) rc = tune_segments(E, reps, Ls, Lf, Ld);
        if (hipStreamSynchronize(E->stream) != hipSuccess && !rc) rc = E->fail(SWM_E_HIP, "stream synchronize");
        for (const Kept &k : kept) {
            if ((hipMemcpy(k.ctx->qs.base, k.buf, k.n, hipMemcpyDeviceToDevice) != hipSuccess ||
                 hipMemcpy(k.ctx->ks.base, k.buf + k.n, k.n, hipMemcpyDeviceToDevice) != hipSuccess) && !rc)
                rc = E->fail(SWM_E_HIP, "kept fields' restore");
            k.ctx->advcor_pending = k.pending;
            (void)hipFree(k.buf);
        }
        return rc;
    }

    static int tune_segments(swm_ens *E, int reps, int *Ls, int *Lf, int *Ld) {
        static const int cand[] = {4, 6, 8, 12, 16, 24, 32, 48};
        for (int cls = 0; cls < 3; ++cls) {
            double best = 1e300;
            int bestL = 0;
            for (int L : cand) {
                for (swm_ctx *ctx : E->m) (cls == 0 ? ctx->L : cls == 1 ? ctx->Lf : ctx->Ld) = (cls == 0 ? std::max(L, 4) : std::max(L, 8));
                double ms = 0;
                int rc = time_step(E, reps, &ms);
                if (rc) return rc;
                if (ms < best) { best = ms; bestL = cls == 0 ? std::max(L, 4) : std::max(L, 8); }
            }
            for (swm_ctx *ctx : E->m) (cls == 0 ? ctx->L : cls == 1 ? ctx->Lf : ctx->Ld) = bestL;
        }
        *Ls = E->m[0]->L;
        *Lf = E->m[0]->Lf;
        *Ld = E->m[0]->Ld;
        E->buf[0].valid = E->buf[1].valid = false;     // plans (and the graph's launch sizes) changed
        E->graph_ok = false;
        return SWM_OK;
    }
};

// the table of entry points of bundle F (the adjoint's only where ADJ)
template <class F, bool ADJ> struct Entry {
    using Q = typename F::Q;
    using ST = typename F::ST;
    static void segment_lengths(swm_ctx *ctx) {
        if (ctx->c.bc == SWM_PERIODIC) Launch<F>::template segment_lengths<true>(ctx);
        else Launch<F>::template segment_lengths<false>(ctx);
    }
    static int scale_static(swm_ctx *ctx) {
        for (Field *g : {&ctx->Fx, &ctx->Fy, &ctx->fq})
            hipLaunchKernelGGL((scale_kernel<ST>), dim3(1024), dim3(256), 0, ctx->stream, (ST *)g->base, ctx->elems,
                               (float)ctx->sc);
        HIPCHK(hipGetLastError());
        return SWM_OK;
    }
    static int upload_fix(swm_ctx *ctx, int pu, int pe) {
        const swm_config &c = ctx->c;
        if (c.scale_log2 != 0)
            for (Field *g : {&ctx->u[pu], &ctx->v[pu]})
                hipLaunchKernelGGL((scale_kernel<Q>), dim3(1024), dim3(256), 0, ctx->stream, (Q *)g->base, ctx->elems,
                                   (float)ctx->sc);
        if (c.compensated)                       // a new state: no compensation carried over
            for (Field *g : {&ctx->cu, &ctx->cv, &ctx->ce})
                hipLaunchKernelGGL(zero_words_kernel, dim3(1024), dim3(256), 0, ctx->stream, (uint32_t *)g->base,
                                   ctx->elems * ctx->es / 4);
        HIPCHK(hipGetLastError());
        return Launch<F>::ghosts(ctx, pu, pe, ctx->rsst);
    }
    static int stages(swm_ctx *ctx, int32_t n, int64_t t0) { return Launch<F>::stages(ctx, n, t0, &ctx->rk0, &ctx->rk1, true); }
    static int reduce_launch(swm_ctx *ctx, int pu, int pe, double *part) {
        const swm_config &c = ctx->c;
        reduce_kernel<Q, ST><<<c.ny, 256, 0, ctx->stream>>>(ctx->at<Q>(ctx->u[pu]), ctx->at<Q>(ctx->v[pu]), ctx->at<Q>(ctx->e[pe]),
                                                            ctx->at<ST>(ctx->H), c.nx, c.ny, ctx->P, ctx->ep, c.g, ctx->sci,
                                                            part);
        HIPCHK(hipGetLastError());
        return SWM_OK;
    }
    static int ghost(swm_ctx *ctx) { return Launch<F>::ghosts(ctx, ctx->pu, ctx->pe, ctx->rsst); }
    static int dissipation(swm_ctx *ctx) {
        const int out = pick({ctx->pu});
        const int rc = Launch<F>::dissipation(ctx, Slots{ctx->pu, ctx->pe}, out);
        if (!rc) ctx->pu = out;
        return rc;
    }
    static int tracer(swm_ctx *ctx, int64_t i) { return Launch<F>::tracer(ctx, i, ctx->pu); }
    // Forcing{T} and f on the device: 2-D fields in the padded layout, the x-invariant
    // ones (winds, f) as one value per row, as swm_set_static stores them
    static int setup_static(swm_ctx *ctx, const swm_setup *st) {
        const swm_config &c = ctx->c;
        SetupArgs a{};
        a.topo = st->topography; a.wind_x = st->wind_x; a.wind_y = st->wind_y;
        a.Lx = st->Lx; a.Ly = st->Ly; a.D = st->Delta; a.H = st->H; a.th = st->topo_height; a.w = st->topo_width;
        a.Fx0 = st->Fx0; a.Fy0 = st->Fy0; a.rho = st->rho; a.f0 = st->f0; a.beta = st->beta;
        a.eta_refh = st->eta_refh; a.eta_refw = st->eta_refw; a.A = st->A; a.phik = st->phik; a.wk = st->wk;
        a.R = st->R; a.phi = st->phi;
        a.nx = c.nx; a.ny = c.ny; a.ep = ctx->ep;
        auto fill = [&](int field, Field &f, bool rows, int m, int n, int x0, int y0) {
            a.field = field;
            if (rows) {
                as_rows(f);
                a.p = f.p0 + (ptrdiff_t)y0 * f.es;
                a.sx = 0; a.sy = 1; a.m = 1;
            } else {
                a.p = elem(ctx, f, x0, y0);
                a.sx = 1; a.sy = ctx->P; a.m = m;
            }
            a.n = n;
            hipLaunchKernelGGL((setup_kernel<ST>), dim3((a.m + 255) / 256, n), dim3(256), 0, ctx->stream, a);
        };
        fill(SET_H, ctx->H, false, c.nx + 2, c.ny + 2, 0, 0);
        fill(SET_FX, ctx->Fx, true, ctx->nux, ctx->nuy, 1 - ctx->ep, 1);
        fill(SET_FY, ctx->Fy, true, ctx->nvx, ctx->nvy, 1, 1);
        fill(SET_FQ, ctx->fq, true, c.nx + 1, c.ny + 1, 0, 0);
        ctx->Fx_sx = ctx->Fy_sx = ctx->fq_sx = 0;
        ctx->Fx_sy = ctx->Fy_sy = ctx->fq_sy = 1;
        if (c.continuity == SWM_CONT_RELAX) fill(SET_ETA_REF, ctx->eta_ref, false, c.nx, c.ny, 1, 1);
        if (c.continuity == SWM_CONT_FORCING) fill(SET_FETA, ctx->Feta, false, c.nx, c.ny, 1, 1);
        if (c.dynamics == SWM_LINEAR) {
            fill(SET_FU, ctx->f_u, true, ctx->nux, ctx->nuy, 1 - ctx->ep, 1);
            fill(SET_FV, ctx->f_v, true, ctx->nvx, ctx->nvy, 1, 1);
            ctx->fu_sx = ctx->fv_sx = 0;
            ctx->fu_sy = ctx->fv_sy = 1;
        }
        HIPCHK(hipGetLastError());
        return SWM_OK;
    }
    static FmtOps make() {
        FmtOps o{};
        o.setup_static = &setup_static;
        o.ghost = &ghost;
        o.dissipation = &dissipation;
        o.tracer = &tracer;
        o.segment_lengths = &segment_lengths;
        o.scale_static = &scale_static;
        o.upload_fix = &upload_fix;
        o.run_steps = &run_steps<F>;
        o.stages = &stages;
        o.time_kernel = &Launch<F>::time_kernel;
        o.reduce_launch = &reduce_launch;
        o.ens_tune = &Ens<F>::tune;
        o.ens_step = &Ens<F>::step;
        if constexpr (ADJ) {
            o.adj_forward = &Adj<F>::forward;
            o.adj_pullback = &Adj<F>::pullback;
            o.adj_time = &Adj<F>::time;
            o.adj_download_fx = &Adj<F>::download_fx;
        }
        return o;
    }
};

}  // namespace swmh
