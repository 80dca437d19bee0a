// k_stage.h -- right-hand side kernels of one Runge-Kutta stage (gfx950).
#pragma once

#include "swm_common.h"

namespace swm {

// Stage combinations of time_integration.jl:53-196, applied pointwise to the tendency d:
//   RK_FIRST   o1 = w + b·d (u1),  o2 = w + a·d (u0)            caxb!/axb!, stage 1 (u0 == u)
//   RK_MID     o1 = pr + b·d,      o2 = ac + a·d                 caxb!(u1,u,...), axb!(u0,...)
//   RK_LAST    o2 = ac + a·d  (+ ghosts of u0, :198)
//   SSP2       o1 = w + c·d                                      axb!(u1,Δt_Δs,du)  :95
//   SSP2_LAST  o1 = w + c·d,  o2 = x·pr + y·o1                   cxayb!(u0,a,u,b,u1) :108
//   SSP3       o1 = w + c·d                                      axb!(u1,Δt_Δn,du)  :132
//   SSP3_KN    o1 = x·w + y·ac + z·d                             dxaybzc!  :129
//   S4         t = w + c·d,  o1 = h·(w + t)                      caxb!(u0,u1,Δt_Δ,du), cxab!  :172-175
//   S4_3       ... and o2 = x·pr + y·o1                          cxayb!(u1,2/3,u,1/3,u1)  :187
//   RKC_FIRST  o1 = w + b·d,  o2 = a·d                 compensated RK: o2 holds Σ a_k·d_k
//   RKC_LAST   Σ = ac + a·d, then o2 = pr + Σ with Kahan's compensated summation (compensation cc)
// w is the stage input at the point, pr the Prog value, ac the u0 value; all in Tprog.
enum { M_RK_FIRST = 0, M_RK_MID, M_RK_LAST, M_SSP2, M_SSP2_LAST, M_SSP3, M_SSP3_KN, M_S4, M_S4_3, M_RKC_FIRST,
       M_RKC_LAST };

template <int M> struct Mode {
    static constexpr bool rk = M <= M_RK_LAST || M == M_RKC_FIRST || M == M_RKC_LAST;   // rhs! and continuity! in one kernel
    static constexpr bool first = M == M_RK_FIRST || M == M_RKC_FIRST;
    static constexpr bool comp = M == M_RKC_LAST;           // reads and writes the compensation
    static constexpr bool need_pr = M == M_RK_MID || M == M_SSP2_LAST || M == M_S4_3 || M == M_RKC_LAST;
    static constexpr bool need_ac = M == M_RK_MID || M == M_RK_LAST || M == M_SSP3_KN || M == M_RKC_LAST;
    static constexpr bool out1 = M != M_RK_LAST && M != M_RKC_LAST;
    static constexpr bool out2 = M == M_RK_FIRST || M == M_RK_MID || M == M_RK_LAST || M == M_SSP2_LAST || M == M_S4_3 ||
                                 M == M_RKC_FIRST || M == M_RKC_LAST;
    static constexpr bool ghost2 = M != M_RK_FIRST && M != M_RK_MID && M != M_RKC_FIRST;   // u0 between RK stages: interior only
};

template <class Q> struct Coef {
    Q a, b, c, x, y, z, h;
};

template <int M, class Q>
__device__ __forceinline__ void combine(Q w, Q pr, Q ac, Q d, const Coef<Q> &k, Q &o1, Q &o2, Q &cc) {
    if constexpr (M == M_RK_FIRST) { o1 = w + k.b * d; o2 = w + k.a * d; }
    else if constexpr (M == M_RKC_FIRST) { o1 = w + k.b * d; o2 = k.a * d; }
    else if constexpr (M == M_RKC_LAST) {
        const Q s = ac + k.a * d;
        const Q y = s - cc;
        const Q t = pr + y;
        cc = (t - pr) - y;
        o2 = t;
    }
    else if constexpr (M == M_RK_MID) { o1 = pr + k.b * d; o2 = ac + k.a * d; }
    else if constexpr (M == M_RK_LAST) { o2 = ac + k.a * d; }
    else if constexpr (M == M_SSP2 || M == M_SSP3) { o1 = w + k.c * d; }
    else if constexpr (M == M_SSP2_LAST) { o1 = w + k.c * d; o2 = k.x * pr + k.y * o1; }
    else if constexpr (M == M_SSP3_KN) { o1 = k.x * w + k.y * ac + k.z * d; }
    else if constexpr (M == M_S4) { const Q t = w + k.c * d; o1 = k.h * (w + t); }
    else { const Q t = w + k.c * d; o1 = k.h * (w + t); o2 = k.x * pr + k.y * o1; }
}

// Field formats: Prog, u0 and the second output live in Tprog (Q); the static fields and the
// stage input u1 between RK stages in T (ST: the reference converts u1 to T before rhs!,
// time_integration.jl:62-64, and uses it for nothing else).  The first stage reads Prog.
template <class F> struct StageArgs {
    using R = typename F::R;
    using Q = typename F::Q;
    ArenaRef ar;                      // the context's fields; the members below are origins in it
    uint32_t in_u, in_v, in_e;        // stage input (u1 of the reference; Prog at stage 1)
    uint32_t pr_u, pr_v, pr_e;        // Prog
    uint32_t ac_u, ac_v, ac_e;        // u0
    uint32_t o1_u, o1_v, o1_e;        // first output (u1 / u1')
    uint32_t o2_u, o2_v, o2_e;        // second output (u0 / u1'')
    uint32_t qend;                    // periodic: q at corner x=nx per row (AHβ/AHγ wrap), 1-D
    uint32_t qs, ks;                  // nstep_advcor > 0: q (corners) and KEu+KEv (T points) kept, 2-D T
    StaticRefs st;
    uint32_t c_u, c_v, c_e;           // compensations of the compensated RK modes (Tprog)
    Coef<Q> k;                        // stage coefficients (first stage of a fused pair)
    Coef<Q> k2;                       // second stage of a fused pair
    // T scalars: k12 = T(1/12)/s, g·s and the ½ of the Bernoulli potential /s for velocities
    // carried scaled by s (= 1 unscaled)
    R Fxt, Fyt, Fet, g, gamma, k12, halfp;
    Q oma, sinv;                      // one_minus_α, 1/s (Tprog)
    int cont, cm, scaled;             // continuity variant, Tcomm, s != 1
    int nx, ny, P, L, nstrips, nseg;
    // fused pairs: strip s of the launch is strip s + sfirst of the field, plus sjump from
    // strip sjump_at on (a launch over the strips at the periodic seam skips the others)
    int sfirst = 0, sjump_at = 1 << 30, sjump = 0;
    uint32_t tp_u = 0, tp_v = 0, tp_e = 0;   // fused pairs with ADV_TAPE: where the first stage's output is kept
};


template <class F, int M> using StageIn = std::conditional_t<Mode<M>::first, typename F::Q, typename F::ST>;

// dη at T-point (x, y) from the flux divergence, continuity.jl:3-68; cs is the static of
// the variant at the point (η_ref for the relaxation, Fη for the forcing).  The divergence
// of fluxes of velocities carried scaled by s is divided by s (sinv = 1/s, exactly 1 when
// unscaled, which changes no value).  CS: the variant may have a static term (cont is then
// read at run time); without it the code has no branch.
template <bool CS, class Q, class R>
__device__ __forceinline__ Q continuity_rhs(R dUdx, R dVdy, R e, int cont, R cs, R gamma, R Fet, Q sinv) {
    Q d = -(toQ<Q>(dUdx) + toQ<Q>(dVdy));
    d = d * sinv;
    if constexpr (CS) {
        if (cont == CONT_RELAX) d = d + toQ<Q>(gamma * (cs - e));
        else if (cont == CONT_FORCING) d = d + toQ<Q>(Fet * cs);
    }
    return d;
}

// The static operands a march takes per row, loaded a row ahead of their use like the
// fields: f_q at corner row y, Fx at u-row y, Fy at v-row y-1, the continuity static at
// T-row y and (periodic) q at the corner (nx, y).  A load that is waited for in the row that
// issues it would also wait for every load issued before it (the counter is in order).
// The linear dynamics read f on the u-grid at row y and on the v-grid at row y-1 instead of
// f_q (fu, fv; column offsets vfu, vfv), and no q column.  With ADV_KEEP the kept q at the
// corner (x, y) and KEu+KEv at the T point (x, y) are read (qs, ks), and the periodic wrap
// column from the kept q at x = nx.
template <class R> struct RowStatics {
    R fq, Fx, Fy, cs, qend, fu, fv, qs, ks;
};

// q at the wrap column of row y: a T value, held as it is (float for T's values with float
// arithmetic, where it is an unrounded result)
template <class F> __device__ __forceinline__ typename F::R ld_qend(const Arena &ar, uint32_t o, int y) {
    using R = typename F::R;
    const Fld<QES<F>> q{&ar, o};
    const auto v = q.ld(0, (uint32_t)(y + OY) * (uint32_t)sizeof(QES<F>));
    if constexpr (is_ar<R>::value) return R::raw(v);
    else return dec<R, QES<F>>(v);
}

template <class F, bool PER, bool CS, int ADV = ADV_AH>
__device__ __forceinline__ RowStatics<typename F::R> row_statics(const StageArgs<F> &A, const Arena &ar, uint32_t vfq,
                                                                 uint32_t vFx, uint32_t vFy, uint32_t vxT, int y,
                                                                 uint32_t vfu = 0, uint32_t vfv = 0) {
    using R = typename F::R;
    using ST = typename F::ST;
    constexpr int SCH = ADV & 3;
    constexpr bool KEEP = (ADV & ADV_KEEP) != 0;
    const int P = A.P;
    RowStatics<R> s;
    s.fq = (SCH == ADV_LIN || KEEP) ? R(0) : Fld<ST>{&ar, A.st.fq}.template ldr<R>(vfq, srowb<ST>(y, P, A.st.fq_2d));
    s.Fx = Fld<ST>{&ar, A.st.Fx}.template ldr<R>(vFx, srowb<ST>(y, P, A.st.Fx_2d));
    s.Fy = Fld<ST>{&ar, A.st.Fy}.template ldr<R>(vFy, srowb<ST>(y - 1, P, A.st.Fy_2d));
    s.cs = (!CS || A.cont == CONT_ITSELF)
               ? R(0)
               : Fld<ST>{&ar, A.cont == CONT_RELAX ? A.st.eta_ref : A.st.Feta}.template ldr<R>(vxT, rowb<ST>(y, P));
    s.qend = !(PER && SCH == ADV_AH) ? R(0)
             : KEEP ? Fld<ST>{&ar, A.qs}.template ldr<R>(colb<ST>(A.nx), rowb<ST>(y, P))
                    : ld_qend<F>(ar, A.qend, y);
    s.fu = SCH == ADV_LIN ? Fld<ST>{&ar, A.st.fu}.template ldr<R>(vfu, srowb<ST>(y, P, A.st.fu_2d)) : R(0);
    s.fv = SCH == ADV_LIN ? Fld<ST>{&ar, A.st.fv}.template ldr<R>(vfv, srowb<ST>(y - 1, P, A.st.fv_2d)) : R(0);
    s.qs = KEEP ? Fld<ST>{&ar, A.qs}.template ldr<R>(vxT, rowb<ST>(y, P)) : R(0);
    s.ks = KEEP ? Fld<ST>{&ar, A.ks}.template ldr<R>(vxT, rowb<ST>(y, P)) : R(0);
    return s;
}

// ------------------------------------------------------------------ right-hand side of one row
//
// rhs_nonlinear!  rhs.jl:24-51 in the y-march of one wavefront: push() takes the stage
// input at row r (u, v, η and H at row r, f_q at corner row r-1) and leaves the terms of
// du and dη at row r-1 and of dv at row r-2 (qhv, the p difference ... ), with the rows
// each stencil needs kept as history.  Arithmetic in T, term by term as the reference.
// ADV_SAD: Sadourny's enstrophy-conserving scheme (PV_advection.jl:31-59) in place of
// Arakawa-Hsu's; then k12 carries 1/2 (the weight of Iy!(q_u,q) and Ix!(q_v,q), rhs.jl:117-120)
// and no column wraps (q_u and q_v read no q[end,·]).  ADV_LIN: rhs_linear! (rhs.jl:55-86),
// the Coriolis terms f·v, f·u of the bilinearly interpolated velocities and the gradient of
// g·η; it forms no mass fluxes, so the divergence it leaves for continuity! is zero (the
// reference's U, V stay as allocated under the classic RK scheme, continuity.jl:83-85).
template <class R, bool PER, int ADV = ADV_AH, bool QOWN = false> struct RhsPipe {
    static constexpr int SCH = ADV & 3;
    static constexpr bool KEEP = (ADV & ADV_KEEP) != 0;
    R u_1 = 0, u_2 = 0, v_1 = 0, v_2 = 0, e_1 = 0, h_1 = 0, hs_1 = 0;
    R U_1 = 0, U_2 = 0, V_2 = 0;
    R q_2 = 0, qaU_2 = 0, qb_2 = 0, p_2 = 0, qend_2 = 0;
    // results of the last push
    R qhv, px;            // row r-1: Arakawa-Hsu qhv and p(x+1) - p(x)  (momentum_u!)
    R dUdx, dVdy, e_r1;   // row r-1: flux divergence terms and η  (continuity!)
    R qhu, py;            // row r-2: Arakawa-Hsu qhu and p(y+1) - p(y)  (momentum_v!)

    // Every x-neighbour is taken where it is consumed, as the shifted operand of one add,
    // subtract or multiply (a lane move the compiler folds into that instruction), and the
    // products of two shifted factors are formed unshifted and moved once:
    // q(x-1)·U(x-1) is the shift of q·U.  The values are those of the formulas as written.
    // qend_1: q at the corner (nx, r-1), read by the wrapped first column when qfix (periodic x = 0)
    // fu_1, fv_2: f on the u-grid at row r-1 and on the v-grid at row r-2 (ADV_LIN); qs_1,
    // ks_1: the kept q and KEu+KEv at row r-1 (ADV_KEEP)
    __device__ __forceinline__ void push(R u_0, R v_0, R e_0, R H_0, R fq_1, R qend_1, bool qfix, R half, R quarter,
                                         R k12, R g, R halfp, R fu_1 = 0, R fv_2 = 0, R qs_1 = 0, R ks_1 = 0) {
        if constexpr (SCH == ADV_LIN) {
            // ∂x!(dpdx,g*η), ∂y!(dpdy,g*η) ; Ixy!(v_u,v), fv! at row r-1 ; Ixy!(u_v,u), fu! at row r-2
            const R p_1 = g * e_1;
            qhv = fu_1 * (quarter * (v_2 + sh_xp(v_2)) + quarter * (v_1 + sh_xp(v_1)));
            qhu = fv_2 * (quarter * (sh_xm(u_2) + u_2) + quarter * (sh_xm(u_1) + u_1));
            px = sh_xp(p_1) - p_1;
            py = p_1 - p_2;
            dUdx = R(0);
            dVdy = R(0);
            e_r1 = e_1;
            u_2 = u_1; u_1 = u_0; v_2 = v_1; v_1 = v_0; e_1 = e_0; p_2 = p_1;
            return;
        }
        front(u_0, v_0, e_0, H_0, fq_1, half, quarter, g, halfp, qs_1, ks_1);
        back(qend_1, qfix, quarter, k12, g);
    }

    // The first part of a push: thickness, fluxes, q and the Bernoulli potential of the row
    // (n_q1 is q at the corner row r-1); back() forms the PV fluxes and the outputs.  Split so
    // that two marches can exchange q between them (the wrap's q[end,·]).
    R n_u0, n_v0, n_e0, n_h0, n_hs0, n_U0, n_V1, n_q1, n_p1;
    __device__ __forceinline__ void front(R u_0, R v_0, R e_0, R H_0, R fq_1, R half, R quarter, R g, R halfp,
                                          R qs_1 = 0, R ks_1 = 0) {
        // thickness!  rhs.jl:131-139 ; Ix!(h_u)  interpolations.jl:3-14 ; Uflux!  rhs.jl:277-291
        n_u0 = u_0; n_v0 = v_0; n_e0 = e_0;
        const R h_0 = e_0 + H_0;
        const R hs_0 = sh_xp(h_0) + h_0;                   // h(x+1) + h, read by Ix!(h_u) and Ixy!(h_q)
        const R U_0 = u_0 * (half * hs_0);
        // row r-1: Iy!(h_v), Vflux!  rhs.jl:294-304
        const R V_1 = v_1 * (half * (h_0 + h_1));
        R q_1, ke_1;
        if constexpr (KEEP) {
            q_1 = qs_1;
            ke_1 = ks_1;
        } else {
            // Ixy!(h_q,h)  interpolations.jl:33-45
            const R hq_1 = quarter * hs_1 + quarter * hs_0;
            // ∂x!(dvdx,v), ∂y!(dudy,u)  gradients.jl:3-25 ; PV!  PV_advection.jl:2-18
            const R dvdx_1 = sh_xp(v_1) - v_1;
            const R dudy_1 = u_0 - u_1;
            q_1 = (fq_1 + dvdx_1 - dudy_1) / hq_1;
            // speed!, Ix!(KEu), Iy!(KEv)  rhs.jl:107-109
            const R uu_1 = u_1 * u_1;
            const R KEu_1 = half * (uu_1 + sh_xm(uu_1));
            const R KEv_1 = half * (v_1 * v_1 + v_2 * v_2);
            ke_1 = KEu_1 + KEv_1;
        }
        // bernoulli!  rhs.jl:162-180
        const R p_1 = halfp * ke_1 + g * e_1;
        n_h0 = h_0; n_hs0 = hs_0; n_U0 = U_0; n_V1 = V_1; n_q1 = q_1; n_p1 = p_1;
    }
    __device__ __forceinline__ void back(R qend_1, bool qfix, R quarter, R k12, R g) {
        const R u_0 = n_u0, v_0 = n_v0, e_0 = n_e0, h_0 = n_h0, hs_0 = n_hs0, U_0 = n_U0, V_1 = n_V1, q_1 = n_q1,
                p_1 = n_p1;
        if constexpr (SCH == ADV_SAD) {
            // qhv = q_u·V_u at row r-1 (q_u = Iy!(q), V_u = Ixy!(V)) and qhu = q_v·U_v at row
            // r-2 (q_v = Ix!(q), U_v = Ixy!(U)); lane x holds the q corner x, the V point
            // x-1/2 and the U point x
            const R qxm_2 = sh_xm(q_2);
            qhv = (k12 * (q_1 + q_2)) * (quarter * (V_2 + sh_xp(V_2)) + quarter * (V_1 + sh_xp(V_1)));
            qhu = (k12 * (q_2 + qxm_2)) * (quarter * (sh_xm(U_2) + U_2) + quarter * (sh_xm(U_1) + U_1));
            px = sh_xp(p_1) - p_1;
            dUdx = U_1 - sh_xm(U_1);
            dVdy = V_1 - V_2;
            e_r1 = e_1;
            py = p_1 - p_2;
            u_1 = u_0; v_2 = v_1; v_1 = v_0; e_1 = e_0; h_1 = h_0; hs_1 = hs_0;
            U_2 = U_1; U_1 = U_0; V_2 = V_1;
            q_2 = q_1; p_2 = p_1;
            return;
        }
        // AHα!..AHδ!  PV_advection.jl:88-146 at row r-1 (β, γ wrap to q[end,·] at x=0).
        // The x-1 neighbours are moved before the per-lane choice of the wrapped column: a
        // lane move inside a branch taken by some lanes only would read 0 from the others.
        // The wrapped β, γ of the first column are read by qhv at the u column 0 only; qhu
        // reads β, γ at the v columns 1..nx, unwrapped (the v column nx too, whose lane also
        // stands for x = 0 when QOWN).  QOWN: q[end,·] is the lane's own q (a march whose
        // lanes wrap around the periodic seam, where q at the corner nx equals q at 0).
        const R s_1 = q_2 + q_1;
        const R qxm_1 = sh_xm(q_1), qxm_2 = sh_xm(q_2);
        const R qa_1 = k12 * (s_1 + sh_xp(q_1));
        const R qbn_1 = k12 * (s_1 + qxm_1);
        // the wrapped first column of AHγ! sums (q[1,j] + q[1,j+1]) + q[end,j] (:126), the
        // other columns (q[i,j] + q[i-1,j]) + q[i,j+1] (:128)
        const R qgn_1 = k12 * ((q_2 + qxm_2) + q_1);
        const R qe_1 = QOWN ? q_1 : qend_1, qe_2 = QOWN ? q_2 : qend_2;
        const R qb_1 = (PER && qfix) ? k12 * (s_1 + qe_1) : qbn_1;
        const R qg_1 = (PER && qfix) ? k12 * (s_1 + qe_2) : qgn_1;
        const R qd_1 = k12 * (s_1 + sh_xp(q_2));
        // PV_ArakawaHsu! qhv  PV_advection.jl:73 and ∂x!(dpdx)  at row r-1
        qhv = qa_1 * sh_xp(V_1) + qb_1 * V_1 + qg_1 * V_2 + qd_1 * sh_xp(V_2);
        px = sh_xp(p_1) - p_1;
        // ∂x!(dUdx,U), ∂y!(dVdy,V)  continuity.jl:52-68 at row r-1
        dUdx = U_1 - sh_xm(U_1);
        dVdy = V_1 - V_2;
        e_r1 = e_1;
        // PV_ArakawaHsu! qhu  PV_advection.jl:81 and ∂y!(dpdy)  at row r-2:
        // qα(x-1)·U(x-1) of row r-2 and qδ(x-1)·U(x-1) of row r-1
        const R qaU_1 = qa_1 * U_1;
        qhu = sh_xm(qaU_2) + qb_2 * U_2 + qgn_1 * U_1 + sh_xm(qd_1 * U_1);
        py = p_1 - p_2;
        // advance the row history
        u_1 = u_0; v_2 = v_1; v_1 = v_0; e_1 = e_0; h_1 = h_0; hs_1 = hs_0;
        U_2 = U_1; U_1 = U_0; V_2 = V_1;
        q_2 = q_1; qaU_2 = qaU_1; qb_2 = qbn_1; p_2 = p_1;
        qend_2 = qend_1;
    }
};

// ------------------------------------------------------------------ stage kernel (K1)
//
// rhs_nonlinear!  rhs.jl:24-51 (UVfluxes!, advection_coriolis!, PV_ArakawaHsu!, bernoulli!,
// ∂x/∂y of p, momentum_u!/momentum_v!), and for the classic RK scheme also continuity!
// (continuity.jl:72-98) on the same stage input, followed by the stage combination and the
// ghost copies of what the next stage reads.  At input row r the wave forms U(r), then the
// row-(r-1) quantities h_v, V, h_q, q, KE, p and the Arakawa-Hsu weights, and emits du (and
// dη) at row r-1 and dv at row r-2.
template <class F, bool PER, int M, int ADV = ADV_AH>
__device__ __forceinline__ void stage_body(const StageArgs<F> &A, int w) {
    using MI = Mode<M>;
    using R = typename F::R;
    using Q = typename F::Q;
    using ST = typename F::ST;
    using SI = StageIn<F, M>;
    using VI = typename Sto<SI>::V;
    const WaveTile t = tile_of<HALO>(w, A.nstrips, A.nseg, A.L, A.ny);
    if (!t.live) return;
    const int x = t.x, P = A.P, nx = A.nx, ny = A.ny;
    const bool lane_ok = t.lane >= HALO && t.lane < WAVE - HALO;
    const bool u_col = lane_ok && (PER ? (x >= 0 && x < nx) : (x >= 1 && x <= nx - 1));
    const bool v_col = lane_ok && x >= 1 && x <= nx;
    const bool e_col = v_col;
    const bool qfix = PER && x == 0;
    const GhostLanes gl = ghost_lanes<PER>(x, nx);

    const Arena ar(A.ar);
    const Fld<SI> in_u{&ar, A.in_u}, in_v{&ar, A.in_v}, in_e{&ar, A.in_e};
    const Fld<Q> pr_u{&ar, A.pr_u}, pr_v{&ar, A.pr_v}, pr_e{&ar, A.pr_e};
    const Fld<Q> ac_u{&ar, A.ac_u}, ac_v{&ar, A.ac_v}, ac_e{&ar, A.ac_e};
    const Fld<ST> o1_u{&ar, A.o1_u}, o1_v{&ar, A.o1_v}, o1_e{&ar, A.o1_e};
    const Fld<Q> o2_u{&ar, A.o2_u}, o2_v{&ar, A.o2_v}, o2_e{&ar, A.o2_e};
    const Fld<ST> fH{&ar, A.st.H};
    const Fld<Q> cmp_u{&ar, A.c_u}, cmp_v{&ar, A.c_v}, cmp_e{&ar, A.c_e};
    // column offsets and row pitches of the three storage formats (equal when T == Tprog)
    const uint32_t vxI = colb<SI>(x), vxQ = colb<Q>(x), vxT = colb<ST>(x);
    const uint32_t dI = (uint32_t)P * sizeof(SI), dQ = (uint32_t)P * sizeof(Q), dT = (uint32_t)P * sizeof(ST);
    const uint32_t vfq = scolb<ST>(x, A.st.fq_2d), vFx = scolb<ST>(x, A.st.Fx_2d), vFy = scolb<ST>(x, A.st.Fy_2d);
    const uint32_t vfu = scolb<ST>(x, A.st.fu_2d), vfv = scolb<ST>(x, A.st.fv_2d);

    const R half = R(0.5), quarter = R(0.25), k12 = A.k12;

    RhsPipe<R, PER, ADV> rhs;
    // the w operands of the combination: the stage input in its own format at rows r-1 (u,
    // η) and r-2 (v)
    VI wu_1 = 0, wv_1 = 0, wv_2 = 0, we_1 = 0;

    // Loads run one row ahead of their use: the stencil inputs of row r+1 and the
    // pointwise operands (Prog, u0) of the rows emitted next iteration are in flight
    // while row r is computed.
    int r = t.ya - 1;
    const int rlast = t.yb + 1;
    uint32_t sI = rowb<SI>(r, P), sQ = rowb<Q>(r, P), sT = rowb<ST>(r, P);   // byte offsets of row r
    VI nu = in_u.ld(vxI, sI), nv = in_v.ld(vxI, sI), ne = in_e.ld(vxI, sI);
    R nH = fH.template ldr<R>(vxT, sT);
    // The last three input rows (yb-1 .. yb+1) are also the first rows of the segment above,
    // which reads them at its start.  Reading them here at the start too, while that wave
    // does, keeps the second read in L2 instead of sending it to HBM at the end of the march.
    VI tu[3], tv[3], te[3];
    R tH[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const uint32_t ri = rowb<SI>(t.yb - 1 + k, P), rt = rowb<ST>(t.yb - 1 + k, P);
        tu[k] = in_u.ld(vxI, ri); tv[k] = in_v.ld(vxI, ri); te[k] = in_e.ld(vxI, ri);
        tH[k] = fH.template ldr<R>(vxT, rt);
    }
    Q npu = 0, nau = 0, npe = 0, nae = 0, npv = 0, nav = 0;   // operands of rows r-1 (u, η) and r-2 (v)
    RowStatics<R> nst = row_statics<F, PER, true, ADV>(A, ar, vfq, vFx, vFy, vxT, r - 1, vfu, vfv);   // f_q, Fx, η static at r-1, Fy at r-2
    for (; r <= rlast; ++r, sI += dI, sQ += dQ, sT += dT) {
        const VI wu_0 = nu, wv_0 = nv, we_0 = ne;
        const R u_0 = dec<R, SI>(wu_0), v_0 = dec<R, SI>(wv_0), e_0 = dec<R, SI>(we_0), H_0 = nH;
        const Q pu_1 = npu, au_1 = nau, pe_1 = npe, ae_1 = nae, pv_2 = npv, av_2 = nav;
        const RowStatics<R> sa = nst;
        nst = row_statics<F, PER, true, ADV>(A, ar, vfq, vFx, vFy, vxT, r, vfu, vfv);
        if (r < rlast) {
            const int k = r + 1 - (t.yb - 1);                  // uniform
            if (k >= 0) {
                nu = k == 0 ? tu[0] : k == 1 ? tu[1] : tu[2];
                nv = k == 0 ? tv[0] : k == 1 ? tv[1] : tv[2];
                ne = k == 0 ? te[0] : k == 1 ? te[1] : te[2];
                nH = k == 0 ? tH[0] : k == 1 ? tH[1] : tH[2];
            } else {
                nu = in_u.ld(vxI, sI + dI); nv = in_v.ld(vxI, sI + dI); ne = in_e.ld(vxI, sI + dI);
                nH = fH.template ldr<R>(vxT, sT + dT);
            }
        }
        if (r >= t.ya && r < t.yb) {                          // row r of u, η is emitted next iteration
            if (MI::need_pr) { npu = pr_u.ld(vxQ, sQ); if (MI::rk) npe = pr_e.ld(vxQ, sQ); }
            if (MI::need_ac) { nau = ac_u.ld(vxQ, sQ); if (MI::rk) nae = ac_e.ld(vxQ, sQ); }
        }
        if (r - 1 >= t.ya && r - 1 < t.yb) {                  // row r-1 of v is emitted next iteration
            if (MI::need_pr) npv = pr_v.ld(vxQ, sQ - dQ);
            if (MI::need_ac) nav = ac_v.ld(vxQ, sQ - dQ);
        }
        rhs.push(u_0, v_0, e_0, H_0, sa.fq, sa.qend, qfix, half, quarter, k12, A.g, A.halfp, sa.fu, sa.fv, sa.qs, sa.ks);

        // du at row r-1: momentum_u!  rhs.jl:242
        const int yu = r - 1;
        const bool row1 = yu >= t.ya && yu < t.yb;
        if (u_col && row1) {
            const Q du = (toQ<Q>(rhs.qhv) - toQ<Q>(rhs.px)) + toQ<Q>(A.Fxt * sa.Fx);
            Q o1 = 0, o2 = 0, cc = 0;
            if (MI::comp) cc = cmp_u.ld(vxQ, sQ - dQ);
            combine<M>((Q)wu_1, MI::first ? (Q)wu_1 : pu_1, au_1, du, A.k, o1, o2, cc);
            if (MI::comp) cmp_u.st(cc, vxQ, sQ - dQ);
            if (MI::out1) put_u<ST, PER>(o1_u, x, vxT, yu, sT - dT, o1, gl.u, nx, ny, P, A.oma, A.cm);
            if (MI::out2) {
                if (MI::ghost2) put_u<Q, PER>(o2_u, x, vxQ, yu, sQ - dQ, o2, gl.u, nx, ny, P, A.oma, A.cm);
                else o2_u.st(o2, vxQ, sQ - dQ);
            }
        }
        // dη at row r-1: continuity!  continuity.jl:52-98
        if (MI::rk && e_col && row1) {
            const Q deta = continuity_rhs<true, Q>(rhs.dUdx, rhs.dVdy, rhs.e_r1, A.cont, sa.cs, A.gamma, A.Fet, A.sinv);
            Q o1 = 0, o2 = 0, cc = 0;
            if (MI::comp) cc = cmp_e.ld(vxQ, sQ - dQ);
            combine<M>((Q)we_1, MI::first ? (Q)we_1 : pe_1, ae_1, deta, A.k, o1, o2, cc);
            if (MI::comp) cmp_e.st(cc, vxQ, sQ - dQ);
            if (MI::out1) put_e<ST, PER>(o1_e, x, vxT, yu, sT - dT, o1, gl.e, nx, ny, P, A.cm);
            if (MI::out2) {
                if (MI::ghost2) put_e<Q, PER>(o2_e, x, vxQ, yu, sQ - dQ, o2, gl.e, nx, ny, P, A.cm);
                else o2_e.st(o2, vxQ, sQ - dQ);
            }
        }
        // dv at row r-2: momentum_v!  rhs.jl:271
        const int yv = r - 2;
        if (v_col && yv >= t.ya && yv < t.yb && yv <= ny - 1) {
            const Q dv = -(toQ<Q>(rhs.qhu) + toQ<Q>(rhs.py)) + toQ<Q>(A.Fyt * sa.Fy);
            Q o1 = 0, o2 = 0, cc = 0;
            if (MI::comp) cc = cmp_v.ld(vxQ, sQ - 2 * dQ);
            combine<M>((Q)wv_2, MI::first ? (Q)wv_2 : pv_2, av_2, dv, A.k, o1, o2, cc);
            if (MI::comp) cmp_v.st(cc, vxQ, sQ - 2 * dQ);
            if (MI::out1) put_v<ST, PER>(o1_v, x, vxT, yv, sT - 2 * dT, o1, gl.v, nx, P, A.oma, A.cm);
            if (MI::out2) {
                if (MI::ghost2) put_v<Q, PER>(o2_v, x, vxQ, yv, sQ - 2 * dQ, o2, gl.v, nx, P, A.oma, A.cm);
                else o2_v.st(o2, vxQ, sQ - 2 * dQ);
            }
        }
        wu_1 = wu_0; wv_2 = wv_1; wv_1 = wv_0; we_1 = we_0;
    }
}

template <class F, bool PER, int M, int ADV = ADV_AH>
__global__ void __launch_bounds__(256) stage_kernel(StageArgs<F> A) {
    stage_body<F, PER, M, ADV>(A, wave_index());
}

// the same stage for every member of an ensemble in one launch (tile table `start`)
template <class F, bool PER, int M, int ADV = ADV_AH>
__global__ void __launch_bounds__(256) stage_batch(const StageArgs<F> *__restrict__ args, const int *__restrict__ start,
                                                   int nmem) {
    int m, wl;
    if (!batch_locate(start, nmem, wave_index(), m, wl)) return;
    const StageArgs<F> A = args[m];
    stage_body<F, PER, M, ADV>(A, wl);
}

// ------------------------------------------------------------------ two fused RK stages
//
// Two consecutive stages of the classic RK scheme in one march (nonperiodic boundaries):
// stage A (mode MA) computes u1, η1 at row r-1 and v1 at row r-2 from the input row r,
// applies ghost_points!(u1,v1,η1) (ghost_points.jl:47-133, time_integration.jl:57) to them
// in registers, and stage B (mode MB) consumes them two rows later as its input; A's second
// output (u0) waits in registers for B's combination.  Only B's outputs are stored, so the
// intermediate stage costs no HBM traffic.  Strips carry an x-halo of 4 lanes (A is valid on
// lanes 2..61, B on 4..59).  Results are those of the two stage kernels, bit for bit.
// CS: the continuity variant has a static term (relaxation or forcing); the plain variant
// runs without the branches that would split the march's body.
constexpr int HALO2 = 4;
constexpr int WOUT2 = WAVE - 2 * HALO2;

// tile w of a fused pair's launch (strips of WAVE - 2·HALO2 output columns)
template <class F> __device__ __forceinline__ WaveTile fused_tile(const StageArgs<F> &A, int w) {
    WaveTile t = tile_of<HALO2>(w, A.nstrips, A.nseg, A.L, A.ny);
    int s = w % A.nstrips + A.sfirst;
    s += s >= A.sjump_at ? A.sjump : 0;
    t.x = -HALO2 + (WAVE - 2 * HALO2) * s + t.lane;
    return t;
}

template <class F, int MA, int MB, bool CS, bool PER = false, int ADV = ADV_AH, bool LEAN = true>
__device__ __forceinline__ void fused_body(const StageArgs<F> &A, int w) {
    using MIA = Mode<MA>;
    using MIB = Mode<MB>;
    using R = typename F::R;
    using Q = typename F::Q;
    using ST = typename F::ST;
    using SI = StageIn<F, MA>;
    using VI = typename Sto<SI>::V;
    static_assert(MIA::rk && MIB::rk && !MIB::first && MIA::out1, "fused pair of classic RK stages");
    constexpr bool TAPE = (ADV & ADV_TAPE) != 0;
    constexpr int AF = ADV & ~ADV_TAPE;                 // the right-hand side form
    static_assert(!TAPE || !F::mixed, "the tape of a fused pair is kept in the stage input's format");
    const WaveTile t = fused_tile(A, w);
    if (!t.live) return;
    const int x = t.x, P = A.P, nx = A.nx, ny = A.ny, lane = t.lane;
    const bool laneB = lane >= HALO2 && lane < WAVE - HALO2;
    const bool u_col = laneB && (PER ? (x >= 0 && x <= nx - 1) : (x >= 1 && x <= nx - 1));
    const bool v_col = laneB && x >= 1 && x <= nx;
    const bool e_col = v_col;
    const GhostLanes gl = ghost_lanes<PER>(x, nx);
    // Periodic: every lane computes stage A at its column taken around the seam (u columns
    // into 0..nx-1, v and η columns into 1..nx), so the lanes beyond the seam hold the values
    // the reference's ghost copies carry, and stage B reads them across lanes.  This needs the
    // ghost copies to be exact (Tcomm no narrower than the conversion to T), H equal on the
    // two sides of the seam and f_q independent of x, which the host checks.
    const int xu = PER ? ((x % nx) + nx) % nx : x;
    const int xt = PER ? xu + (xu == 0 ? nx : 0) : x;
    const bool qfix = PER && xu == 0;                  // lanes of the u column 0 (AHβ/AHγ wrap)
    // waves whose stage-A lanes (2..61) reach the x-halo apply the ghost copies across lanes
    const bool xedge = __any((lane >= 2 && lane < WAVE - 2) && (x <= 0 || x >= nx + 1)) != 0;
    const bool u_halo = x <= 0 || x >= nx;             // normal-velocity halo of u: never written (0)

    const Arena ar(A.ar);
    const Fld<SI> in_u{&ar, A.in_u}, in_v{&ar, A.in_v}, in_e{&ar, A.in_e};
    const Fld<Q> pr_u{&ar, A.pr_u}, pr_v{&ar, A.pr_v}, pr_e{&ar, A.pr_e};
    const Fld<Q> ac_u{&ar, A.ac_u}, ac_v{&ar, A.ac_v}, ac_e{&ar, A.ac_e};
    const Fld<ST> o1_u{&ar, A.o1_u}, o1_v{&ar, A.o1_v}, o1_e{&ar, A.o1_e};
    const Fld<Q> o2_u{&ar, A.o2_u}, o2_v{&ar, A.o2_v}, o2_e{&ar, A.o2_e};
    const Fld<ST> fH{&ar, A.st.H};
    const Fld<Q> cmp_u{&ar, A.c_u}, cmp_v{&ar, A.c_v}, cmp_e{&ar, A.c_e};
    const Fld<ST> tp_u{&ar, A.tp_u}, tp_v{&ar, A.tp_v}, tp_e{&ar, A.tp_e};
    const uint32_t vxQ = colb<Q>(x), vxT = colb<ST>(x);                 // stores (output lanes)
    const uint32_t vIu = colb<SI>(xu), vIt = colb<SI>(xt), vQu = colb<Q>(xu), vQt = colb<Q>(xt), vTt = colb<ST>(xt);
    const uint32_t vfq = scolb<ST>(xu, A.st.fq_2d), vFx = scolb<ST>(xu, A.st.Fx_2d), vFy = scolb<ST>(xt, A.st.Fy_2d);
    const uint32_t vfu = scolb<ST>(xu, A.st.fu_2d), vfv = scolb<ST>(xt, A.st.fv_2d);
    const R half = R(0.5), quarter = R(0.25), k12 = A.k12;
    const Q oma = A.oma;

    RhsPipe<R, PER, AF, PER> ra, rb;
    VI wu_1 = 0, wv_1 = 0, wv_2 = 0, we_1 = 0;        // stage A's input at rows r-1 (u, η), r-2 (v)
    // LEAN (the default): the loads that would read HBM a second time are avoided at the cost
    // of a few registers; callers short of registers pass false.  A first stage reads Prog as
    // its input, which B then needs three rows later (its Prog operand, u and η at row r-3,
    // v at r-4): kept in registers rather than loaded again, as the load would find it gone
    // from L2.  Own lanes only are stored by B, and there A's input column is B's column
    // (xu = xt = x).
    constexpr bool PRHIST = LEAN && MIA::first && MIB::need_pr;
    VI wu_2 = 0, wu_3 = 0, we_2 = 0, we_3 = 0, wv_3 = 0, wv_4 = 0;
    // stage A's outputs, k iterations old: u1, η1 of row r-1-k, v1 of row r-2-k, u0 / η0 / v0 alike
    Q U1_0 = 0, U1_1 = 0, E1_0 = 0, E1_1 = 0, V1_0 = 0;
    Q U0_0 = 0, U0_1 = 0, U0_2 = 0, E0_0 = 0, E0_1 = 0, E0_2 = 0, V0_0 = 0, V0_1 = 0, V0_2 = 0;
    R H_1 = 0, H_2 = 0;                                // H at rows r-1, r-2

    // A marches input rows ya-2 .. yb+3 (its outputs from row ya-1 on are valid); B takes
    // A's row r-2 at iteration r and emits u, η at row r-3 and v at row r-4.
    int r = t.ya - 2;
    const int rlast = t.yb + 3;
    VI nu = in_u.ld(vIu, rowb<SI>(r, P)), nv = in_v.ld(vIt, rowb<SI>(r, P)), ne = in_e.ld(vIt, rowb<SI>(r, P));
    R nH = fH.template ldr<R>(vTt, rowb<ST>(r, P));
    Q npu = 0, nau = 0, npe = 0, nae = 0, npv = 0, nav = 0;      // A's Prog / u0 operands (rows r-1, r-2)
    Q bpu = 0, bpe = 0, bpv = 0;                                 // B's Prog operands (rows r-3, r-4)
    // A's statics a row ahead (f_q, Fx, η static at row r-1, Fy at r-2); B takes those A had
    // two iterations before (rows r-3, r-4)
    RowStatics<R> nst = row_statics<F, false, CS, AF>(A, ar, vfq, vFx, vFy, vTt, r - 1, vfu, vfv);
    RowStatics<R> st_1 = {0, 0, 0, 0, 0}, st_2 = {0, 0, 0, 0, 0};
    const uint32_t dI = (uint32_t)P * sizeof(SI), dQ = (uint32_t)P * sizeof(Q), dT = (uint32_t)P * sizeof(ST);
    uint32_t sI = rowb<SI>(r, P), sQ = rowb<Q>(r, P), sT = rowb<ST>(r, P);   // byte offsets of row r
    // Two rows per trip: the row history is then in the same registers at the top of every
    // trip, and the compiler has no copies to make for it.  An odd count of rows gets one
    // more at the end, which stores nothing (u, η at row yb+1, v at yb) and reads rows up
    // to yb+4 of the padded fields.
    const int rend = rlast + ((rlast - r + 1) & 1);
    auto row = [&](const int r, const uint32_t sI, const uint32_t sQ, const uint32_t sT) __attribute__((always_inline)) {
        const VI wu_0 = nu, wv_0 = nv, we_0 = ne;
        const R u_0 = dec<R, SI>(wu_0), v_0 = dec<R, SI>(wv_0), e_0 = dec<R, SI>(we_0), H_0 = nH;
        const Q pu_1 = npu, au_1 = nau, pe_1 = npe, ae_1 = nae, pv_2 = npv, av_2 = nav;
        const Q qpu_3 = PRHIST ? (Q)wu_3 : bpu, qpe_3 = PRHIST ? (Q)we_3 : bpe, qpv_4 = PRHIST ? (Q)wv_4 : bpv;
        const RowStatics<R> sa = nst, sb = st_2;
        nst = row_statics<F, false, CS, AF>(A, ar, vfq, vFx, vFy, vTt, r, vfu, vfv);
        {   // the input row r+1, clamped to the last row the march reads
            const int rn = min(r + 1, rlast);
            nu = in_u.ld(vIu, rowb<SI>(rn, P)); nv = in_v.ld(vIt, rowb<SI>(rn, P)); ne = in_e.ld(vIt, rowb<SI>(rn, P));
            nH = fH.template ldr<R>(vTt, rowb<ST>(rn, P));
        }
        {   // A's pointwise operands for the next iteration: rows r (u, η) and r-1 (v)
            if (MIA::need_pr) { npu = pr_u.ld(vQu, sQ); npe = pr_e.ld(vQt, sQ); npv = pr_v.ld(vQt, sQ - dQ); }
            if (MIA::need_ac) {
                // A's u0 reaches B's combination on the segment's rows only (LEAN: the rows of
                // the march beyond it load a row of the segment instead, a cache hit never used)
                const uint32_t su = LEAN ? rowb<Q>(min(max(r, t.ya), t.yb - 1), P) : sQ;
                const uint32_t sv = LEAN ? rowb<Q>(min(max(r - 1, t.ya), t.yb - 1), P) : sQ - dQ;
                nau = ac_u.ld(vQu, su); nae = ac_e.ld(vQt, su); nav = ac_v.ld(vQt, sv);
            }
            // B's Prog operands for the next iteration: rows r-2 (u, η) and r-3 (v)
            if (MIB::need_pr && !PRHIST) {
                bpu = pr_u.ld(vxQ, sQ - 2 * dQ); bpe = pr_e.ld(vxQ, sQ - 2 * dQ); bpv = pr_v.ld(vxQ, sQ - 3 * dQ);
            }
        }

        // ---- stage A at input row r: u1, η1, u0, η0 at row r-1 and v1, v0 at row r-2
        ra.push(u_0, v_0, e_0, H_0, sa.fq, R(0), qfix, half, quarter, k12, A.g, A.halfp, sa.fu, sa.fv, sa.qs, sa.ks);
        Q cu1, cu0, ce1, ce0, cv1, cv0;
        {
            const Q du = (toQ<Q>(ra.qhv) - toQ<Q>(ra.px)) + toQ<Q>(A.Fxt * sa.Fx);
            Q nocc = 0;
            combine<MA>((Q)wu_1, MIA::first ? (Q)wu_1 : pu_1, au_1, du, A.k, cu1, cu0, nocc);
            const Q deta = continuity_rhs<CS, Q>(ra.dUdx, ra.dVdy, ra.e_r1, A.cont, sa.cs, A.gamma, A.Fet, A.sinv);
            combine<MA>((Q)we_1, MIA::first ? (Q)we_1 : pe_1, ae_1, deta, A.k, ce1, ce0, nocc);
            const Q dv = -(toQ<Q>(ra.qhu) + toQ<Q>(ra.py)) + toQ<Q>(A.Fyt * sa.Fy);
            combine<MA>((Q)wv_2, MIA::first ? (Q)wv_2 : pv_2, av_2, dv, A.k, cv1, cv0, nocc);
        }
        if constexpr (TAPE) {
            // A's first outputs as the stage kernel stores them: the rows of this segment,
            // the columns B stores, with the ghost copies
            const int yu = r - 1, yv = r - 2;
            if (yu >= t.ya && yu < t.yb) {
                if (u_col) put_u<ST, PER>(tp_u, x, vxT, yu, sT - dT, cu1, gl.u, nx, ny, P, oma, A.cm);
                if (e_col) put_e<ST, PER>(tp_e, x, vxT, yu, sT - dT, ce1, gl.e, nx, ny, P, A.cm);
            }
            if (v_col && yv >= t.ya && yv < t.yb && yv <= ny - 1)
                put_v<ST, PER>(tp_v, x, vxT, yv, sT - 2 * dT, cv1, gl.v, nx, P, oma, A.cm);
        }
        // ghost_points!(u1, v1, η1) on A's first outputs.  Nonperiodic: u's normal halo is
        // 0, its tangential rows mirror with (1-α); v's normal rows are 0, its columns mirror
        // with (1-α); η copies its neighbour row / column.  Periodic: the columns beyond the
        // seam hold the copies already, u's rows beyond the walls are never written (0).
        // Row 0 (first segment) copies row 1, which A has not made yet: B fixes it when it
        // reads row 0.
        if (!PER && u_halo) cu1 = Q(0);
        if (r - 1 == ny + 1) cu1 = PER ? Q(0) : oma * U1_0;
        if (r - 1 == ny + 2) cu1 = PER ? Q(0) : oma * U1_1;
        if (r - 1 == ny + 1) ce1 = E1_0;
        if (r - 2 <= 0 || r - 2 >= ny) cv1 = Q(0);
        if (!PER && xedge) {
            const Q vp1 = sh_xp(cv1), vm1 = sh_xm(cv1);
            const Q vp3 = __shfl(cv1, lane + 3), vm3 = __shfl(cv1, lane - 3);
            const Q ep1 = sh_xp(ce1), em1 = sh_xm(ce1);
            if (x == 0) { cv1 = oma * vp1; ce1 = ep1; }
            if (x == -1) cv1 = oma * vp3;
            if (x == nx + 1) { cv1 = oma * vm1; ce1 = em1; }
            if (x == nx + 2) cv1 = oma * vm3;
        }
        U1_1 = U1_0; U1_0 = cu1; E1_1 = E1_0; E1_0 = ce1; V1_0 = cv1;
        U0_2 = U0_1; U0_1 = U0_0; U0_0 = cu0; E0_2 = E0_1; E0_1 = E0_0; E0_0 = ce0;
        V0_2 = V0_1; V0_1 = V0_0; V0_0 = cv0;

        // ---- stage B at its input row r-2.  B pushes from the first iteration on, so that its
        // row history is carried in the same registers on every iteration; the rows it takes
        // before ya-1 are A's halo rows, and what they give reaches no stored output (u and η
        // from B's input row ya+1 on, v from ya+2 on, each reading two rows of history).
        const int rbi = r - 2;
        {
            const Q ub = rbi == 0 ? (PER ? Q(0) : oma * U1_0) : U1_1;   // row 0: (1-α)·u1 of row 1 (periodic: 0)
            const Q eb = rbi == 0 ? E1_0 : E1_1;                // row 0: η1 of row 1
            rb.push(toR<R>(ub), toR<R>(V1_0), toR<R>(eb), H_2, sb.fq, R(0), qfix, half, quarter, k12, A.g, A.halfp,
                    sb.fu, sb.fv, sb.qs, sb.ks);
            const int yu = r - 3;
            if (yu >= t.ya && yu < t.yb) {
                if (u_col) {
                    const Q du = (toQ<Q>(rb.qhv) - toQ<Q>(rb.px)) + toQ<Q>(A.Fxt * sb.Fx);
                    Q o1 = 0, o2 = 0, cc = 0;
                    if (MIB::comp) cc = cmp_u.ld(vxQ, sQ - 3 * dQ);
                    combine<MB>(Q(0), qpu_3, U0_2, du, A.k2, o1, o2, cc);
                    if (MIB::comp) cmp_u.st(cc, vxQ, sQ - 3 * dQ);
                    if (MIB::out1) put_u<ST, PER>(o1_u, x, vxT, yu, sT - 3 * dT, o1, gl.u, nx, ny, P, oma, A.cm);
                    if (MIB::ghost2) put_u<Q, PER>(o2_u, x, vxQ, yu, sQ - 3 * dQ, o2, gl.u, nx, ny, P, oma, A.cm);
                    else o2_u.st(o2, vxQ, sQ - 3 * dQ);
                }
                if (e_col) {
                    const Q deta = continuity_rhs<CS, Q>(rb.dUdx, rb.dVdy, rb.e_r1, A.cont, sb.cs, A.gamma, A.Fet, A.sinv);
                    Q o1 = 0, o2 = 0, cc = 0;
                    if (MIB::comp) cc = cmp_e.ld(vxQ, sQ - 3 * dQ);
                    combine<MB>(Q(0), qpe_3, E0_2, deta, A.k2, o1, o2, cc);
                    if (MIB::comp) cmp_e.st(cc, vxQ, sQ - 3 * dQ);
                    if (MIB::out1) put_e<ST, PER>(o1_e, x, vxT, yu, sT - 3 * dT, o1, gl.e, nx, ny, P, A.cm);
                    if (MIB::ghost2) put_e<Q, PER>(o2_e, x, vxQ, yu, sQ - 3 * dQ, o2, gl.e, nx, ny, P, A.cm);
                    else o2_e.st(o2, vxQ, sQ - 3 * dQ);
                }
            }
            const int yv = r - 4;
            if (v_col && yv >= t.ya && yv < t.yb && yv <= ny - 1) {
                const Q dv = -(toQ<Q>(rb.qhu) + toQ<Q>(rb.py)) + toQ<Q>(A.Fyt * sb.Fy);
                Q o1 = 0, o2 = 0, cc = 0;
                if (MIB::comp) cc = cmp_v.ld(vxQ, sQ - 4 * dQ);
                combine<MB>(Q(0), qpv_4, V0_2, dv, A.k2, o1, o2, cc);
                if (MIB::comp) cmp_v.st(cc, vxQ, sQ - 4 * dQ);
                if (MIB::out1) put_v<ST, PER>(o1_v, x, vxT, yv, sT - 4 * dT, o1, gl.v, nx, P, oma, A.cm);
                if (MIB::ghost2) put_v<Q, PER>(o2_v, x, vxQ, yv, sQ - 4 * dQ, o2, gl.v, nx, P, oma, A.cm);
                else o2_v.st(o2, vxQ, sQ - 4 * dQ);
            }
        }
        if constexpr (PRHIST) { wu_3 = wu_2; wu_2 = wu_1; we_3 = we_2; we_2 = we_1; wv_4 = wv_3; wv_3 = wv_2; }
        wu_1 = wu_0; wv_2 = wv_1; wv_1 = wv_0; we_1 = we_0;
        H_2 = H_1; H_1 = H_0;
        st_2 = st_1; st_1 = sa;
    };
    for (; r <= rend; r += 2, sI += 2 * dI, sQ += 2 * dQ, sT += 2 * dT) {
        row(r, sI, sQ, sT);
        row(r + 1, sI + dI, sQ + dQ, sT + dT);
    }
}

// A fused pair on the strips that reach the periodic seam, for the seams the lanes cannot
// simply wrap around: H's ghost columns hold the topography of the halo grid (grid.jl:59),
// not the columns they face, and a Tcomm narrower than T rounds the ghost copies.  Every
// lane runs stage A twice, as its own column x over the stored fields (ghost columns
// included: the reference's view of its interior columns) and as the column xs = x ± nx
// across the seam (the view of the columns the ghost copies come from).  A's outputs at
// the ghost lanes are the second view's, rounded through Tcomm as ghost_points! rounds its
// copies; B runs on those in the first view, and the q[end,·] of its AHβ/AHγ wrap (q at
// the corner nx of B's input) is formed from B's inputs as the second view sees them.  A's
// own q[end,·] is the q the other view forms at the same lane (the corner nx).
template <class F, int MA, int MB, bool CS, bool LEAN = true>
__device__ __forceinline__ void fused_seam_body(const StageArgs<F> &A, int w) {
    using MIA = Mode<MA>;
    using MIB = Mode<MB>;
    using R = typename F::R;
    using Q = typename F::Q;
    using ST = typename F::ST;
    using SI = StageIn<F, MA>;
    using VI = typename Sto<SI>::V;
    static_assert(MIA::rk && MIB::rk && !MIB::first && MIA::out1, "fused pair of classic RK stages");
    const WaveTile t = fused_tile(A, w);
    if (!t.live) return;
    const int x = t.x, P = A.P, nx = A.nx, ny = A.ny, lane = t.lane;
    const bool laneB = lane >= HALO2 && lane < WAVE - HALO2;
    const bool u_col = laneB && x >= 0 && x <= nx - 1;
    const bool v_col = laneB && x >= 1 && x <= nx;
    const bool e_col = v_col;
    const GhostLanes gl = ghost_lanes<true>(x, nx);
    const int xs = x < nx / 2 ? x + nx : x - nx;       // the column across the seam
    // ghost lanes of column x (u: beyond 0..nx-1, v and η: beyond 1..nx); they are the
    // interior lanes of column xs, and the other way round
    const bool ug = x <= -1 || x >= nx, vg = x <= 0 || x >= nx + 1;
    const bool qfix = x == 0, qfixs = xs == 0;        // u column 0: the AHβ/AHγ wrap

    const Arena ar(A.ar);
    const Fld<SI> in_u{&ar, A.in_u}, in_v{&ar, A.in_v}, in_e{&ar, A.in_e};
    const Fld<Q> pr_u{&ar, A.pr_u}, pr_v{&ar, A.pr_v}, pr_e{&ar, A.pr_e};
    const Fld<Q> ac_u{&ar, A.ac_u}, ac_v{&ar, A.ac_v}, ac_e{&ar, A.ac_e};
    const Fld<ST> o1_u{&ar, A.o1_u}, o1_v{&ar, A.o1_v}, o1_e{&ar, A.o1_e};
    const Fld<Q> o2_u{&ar, A.o2_u}, o2_v{&ar, A.o2_v}, o2_e{&ar, A.o2_e};
    const Fld<ST> fH{&ar, A.st.H};
    const Fld<Q> cmp_u{&ar, A.c_u}, cmp_v{&ar, A.c_v}, cmp_e{&ar, A.c_e};
    const uint32_t vxI = colb<SI>(x), vxQ = colb<Q>(x), vxT = colb<ST>(x);
    const uint32_t vsI = colb<SI>(xs), vsQ = colb<Q>(xs), vsT = colb<ST>(xs);
    const uint32_t vfq = scolb<ST>(x, A.st.fq_2d), vFx = scolb<ST>(x, A.st.Fx_2d), vFy = scolb<ST>(x, A.st.Fy_2d);
    const uint32_t sfq = scolb<ST>(xs, A.st.fq_2d), sFx = scolb<ST>(xs, A.st.Fx_2d), sFy = scolb<ST>(xs, A.st.Fy_2d);
    const R half = R(0.5), quarter = R(0.25), k12 = A.k12;
    const int cm = A.cm;

    RhsPipe<R, true> ra, ras, rb;                      // A as column x, A as column xs, B
    VI wu_1 = 0, wv_1 = 0, wv_2 = 0, we_1 = 0, su_1 = 0, sv_1 = 0, sv_2 = 0, se_1 = 0;
    // LEAN: a first stage's input as column x is B's Prog operand three rows later (u, η at
    // row r-3, v at r-4): kept in registers rather than read from HBM a second time
    constexpr bool PRHIST = LEAN && MIA::first && MIB::need_pr;
    VI wu_2 = 0, wu_3 = 0, we_2 = 0, we_3 = 0, wv_3 = 0, wv_4 = 0;
    Q U1_0 = 0, U1_1 = 0, E1_0 = 0, E1_1 = 0, V1_0 = 0;       // B's input as column x sees it
    Q W1_0 = 0, W1_1 = 0, F1_0 = 0, F1_1 = 0, X1_0 = 0;       // ... and as column xs sees it (u, η, v)
    Q U0_0 = 0, U0_1 = 0, U0_2 = 0, E0_0 = 0, E0_1 = 0, E0_2 = 0, V0_0 = 0, V0_1 = 0, V0_2 = 0;
    R H_1 = 0, H_2 = 0, Hs_1 = 0, Hs_2 = 0;
    // B's inputs as the column xs sees them, and its q at the corner nx (lane x = 0)
    R bu_1 = 0, bv_1 = 0, bhs_1 = 0;

    int r = t.ya - 2;
    const int rlast = t.yb + 3;
    VI nu = in_u.ld(vxI, rowb<SI>(r, P)), nv = in_v.ld(vxI, rowb<SI>(r, P)), ne = in_e.ld(vxI, rowb<SI>(r, P));
    VI mu = in_u.ld(vsI, rowb<SI>(r, P)), mv = in_v.ld(vsI, rowb<SI>(r, P)), me = in_e.ld(vsI, rowb<SI>(r, P));
    R nH = fH.template ldr<R>(vxT, rowb<ST>(r, P)), mH = fH.template ldr<R>(vsT, rowb<ST>(r, P));
    Q npu = 0, nau = 0, npe = 0, nae = 0, npv = 0, nav = 0;
    Q mpu = 0, mau = 0, mpe = 0, mae = 0, mpv = 0, mav = 0;
    Q bpu = 0, bpe = 0, bpv = 0;
    RowStatics<R> nst = row_statics<F, false, CS>(A, ar, vfq, vFx, vFy, vxT, r - 1);
    RowStatics<R> mst = row_statics<F, false, CS>(A, ar, sfq, sFx, sFy, vsT, r - 1);
    RowStatics<R> st_1 = {}, st_2 = {}, ss_1 = {}, ss_2 = {};
    const uint32_t dQ = (uint32_t)P * sizeof(Q), dT = (uint32_t)P * sizeof(ST);
    uint32_t sQ = rowb<Q>(r, P), sT = rowb<ST>(r, P);
    for (; r <= rlast; ++r, sQ += dQ, sT += dT) {
        const VI wu_0 = nu, wv_0 = nv, we_0 = ne, su_0 = mu, sv_0 = mv, se_0 = me;
        const R H_0 = nH, Hs_0 = mH;
        const Q pu_1 = npu, au_1 = nau, pe_1 = npe, ae_1 = nae, pv_2 = npv, av_2 = nav;
        const Q qpu_1 = mpu, qau_1 = mau, qpe_1 = mpe, qae_1 = mae, qpv_2 = mpv, qav_2 = mav;
        const Q qpu_3 = PRHIST ? (Q)wu_3 : bpu, qpe_3 = PRHIST ? (Q)we_3 : bpe, qpv_4 = PRHIST ? (Q)wv_4 : bpv;
        const RowStatics<R> sa = nst, sas = mst, sb = st_2, sbs = ss_2;
        nst = row_statics<F, false, CS>(A, ar, vfq, vFx, vFy, vxT, r);
        mst = row_statics<F, false, CS>(A, ar, sfq, sFx, sFy, vsT, r);
        {
            const int rn = min(r + 1, rlast);
            const uint32_t so = rowb<SI>(rn, P), to = rowb<ST>(rn, P);
            nu = in_u.ld(vxI, so); nv = in_v.ld(vxI, so); ne = in_e.ld(vxI, so); nH = fH.template ldr<R>(vxT, to);
            mu = in_u.ld(vsI, so); mv = in_v.ld(vsI, so); me = in_e.ld(vsI, so); mH = fH.template ldr<R>(vsT, to);
        }
        if (MIA::need_pr) {
            npu = pr_u.ld(vxQ, sQ); npe = pr_e.ld(vxQ, sQ); npv = pr_v.ld(vxQ, sQ - dQ);
            mpu = pr_u.ld(vsQ, sQ); mpe = pr_e.ld(vsQ, sQ); mpv = pr_v.ld(vsQ, sQ - dQ);
        }
        if (MIA::need_ac) {       // A's u0 reaches B's combination on the segment's rows only (LEAN)
            const uint32_t su = LEAN ? rowb<Q>(min(max(r, t.ya), t.yb - 1), P) : sQ;
            const uint32_t sv = LEAN ? rowb<Q>(min(max(r - 1, t.ya), t.yb - 1), P) : sQ - dQ;
            nau = ac_u.ld(vxQ, su); nae = ac_e.ld(vxQ, su); nav = ac_v.ld(vxQ, sv);
            mau = ac_u.ld(vsQ, su); mae = ac_e.ld(vsQ, su); mav = ac_v.ld(vsQ, sv);
        }
        if (MIB::need_pr && !PRHIST) {
            bpu = pr_u.ld(vxQ, sQ - 2 * dQ); bpe = pr_e.ld(vxQ, sQ - 2 * dQ); bpv = pr_v.ld(vxQ, sQ - 3 * dQ);
        }

        // ---- stage A at input row r in both views.  The wrap's q[end,·] of one view is q at
        // the corner nx, which the other view forms at the same lane.
        ra.front(dec<R, SI>(wu_0), dec<R, SI>(wv_0), dec<R, SI>(we_0), H_0, sa.fq, half, quarter, A.g, A.halfp);
        ras.front(dec<R, SI>(su_0), dec<R, SI>(sv_0), dec<R, SI>(se_0), Hs_0, sas.fq, half, quarter, A.g, A.halfp);
        {
            const R qa = ra.n_q1, qs = ras.n_q1;
            ra.back(qs, qfix, quarter, k12, A.g);
            ras.back(qa, qfixs, quarter, k12, A.g);
        }
        Q cu1, cu0, ce1, ce0, cv1, cv0, du1, du0, de1, de0, dv1, dv0, nocc = 0;
        {
            const Q du = (toQ<Q>(ra.qhv) - toQ<Q>(ra.px)) + toQ<Q>(A.Fxt * sa.Fx);
            combine<MA>((Q)wu_1, MIA::first ? (Q)wu_1 : pu_1, au_1, du, A.k, cu1, cu0, nocc);
            const Q deta = continuity_rhs<CS, Q>(ra.dUdx, ra.dVdy, ra.e_r1, A.cont, sa.cs, A.gamma, A.Fet, A.sinv);
            combine<MA>((Q)we_1, MIA::first ? (Q)we_1 : pe_1, ae_1, deta, A.k, ce1, ce0, nocc);
            const Q dv = -(toQ<Q>(ra.qhu) + toQ<Q>(ra.py)) + toQ<Q>(A.Fyt * sa.Fy);
            combine<MA>((Q)wv_2, MIA::first ? (Q)wv_2 : pv_2, av_2, dv, A.k, cv1, cv0, nocc);
        }
        {
            const Q du = (toQ<Q>(ras.qhv) - toQ<Q>(ras.px)) + toQ<Q>(A.Fxt * sas.Fx);
            combine<MA>((Q)su_1, MIA::first ? (Q)su_1 : qpu_1, qau_1, du, A.k, du1, du0, nocc);
            const Q deta = continuity_rhs<CS, Q>(ras.dUdx, ras.dVdy, ras.e_r1, A.cont, sas.cs, A.gamma, A.Fet, A.sinv);
            combine<MA>((Q)se_1, MIA::first ? (Q)se_1 : qpe_1, qae_1, deta, A.k, de1, de0, nocc);
            const Q dv = -(toQ<Q>(ras.qhu) + toQ<Q>(ras.py)) + toQ<Q>(A.Fyt * sas.Fy);
            combine<MA>((Q)sv_2, MIA::first ? (Q)sv_2 : qpv_2, qav_2, dv, A.k, dv1, dv0, nocc);
        }
        // A's first outputs as B reads them (ghost_points!): as column x, the ghost lanes hold
        // the copies, through Tcomm, of the columns across the seam; as column xs, the lanes
        // that are ghosts there hold the copies of column x.  u's rows beyond the walls are
        // never written (0), v's wall rows are 0, η's row ny+1 copies row ny.
        Q wu1 = ug ? du1 : comm_round(cu1, cm), wv1 = vg ? dv1 : comm_round(cv1, cm), we1 = vg ? de1 : comm_round(ce1, cm);
        if (ug) cu1 = comm_round(du1, cm);
        if (vg) { cv1 = comm_round(dv1, cm); ce1 = comm_round(de1, cm); }
        if (r - 1 == ny + 1 || r - 1 == ny + 2) { cu1 = Q(0); wu1 = Q(0); }
        if (r - 1 == ny + 1) { ce1 = E1_0; we1 = F1_0; }
        if (r - 2 <= 0 || r - 2 >= ny) { cv1 = Q(0); wv1 = Q(0); }
        U1_1 = U1_0; U1_0 = cu1; E1_1 = E1_0; E1_0 = ce1; V1_0 = cv1;
        W1_1 = W1_0; W1_0 = wu1; F1_1 = F1_0; F1_0 = we1; X1_0 = wv1;
        U0_2 = U0_1; U0_1 = U0_0; U0_0 = cu0; E0_2 = E0_1; E0_1 = E0_0; E0_0 = ce0;
        V0_2 = V0_1; V0_1 = V0_0; V0_0 = cv0;

        // ---- stage B at its input row r-2
        const int rbi = r - 2;
        {
            const Q ub = rbi == 0 ? Q(0) : U1_1;
            const Q eb = rbi == 0 ? E1_0 : E1_1;
            // q at the corner nx of B's input (the wrap's q[end,·]) as the column xs sees it
            {
                const R us = toR<R>(rbi == 0 ? Q(0) : W1_1);
                const R vs = toR<R>(X1_0);
                const R es = toR<R>(rbi == 0 ? F1_0 : F1_1);
                const R hs0 = es + Hs_2;
                const R hss_0 = sh_xp(hs0) + hs0;
                const R hq = quarter * bhs_1 + quarter * hss_0;
                const R dvdx = sh_xp(bv_1) - bv_1;
                const R dudy = us - bu_1;
                const R qe = (sbs.fq + dvdx - dudy) / hq;
                rb.push(toR<R>(ub), toR<R>(V1_0), toR<R>(eb), H_2, sb.fq, qe, qfix, half, quarter, k12, A.g, A.halfp);
                bu_1 = us; bv_1 = vs; bhs_1 = hss_0;
            }
            const int yu = r - 3;
            if (yu >= t.ya && yu < t.yb) {
                if (u_col) {
                    const Q du = (toQ<Q>(rb.qhv) - toQ<Q>(rb.px)) + toQ<Q>(A.Fxt * sb.Fx);
                    Q o1 = 0, o2 = 0, cc = 0;
                    if (MIB::comp) cc = cmp_u.ld(vxQ, sQ - 3 * dQ);
                    combine<MB>(Q(0), qpu_3, U0_2, du, A.k2, o1, o2, cc);
                    if (MIB::comp) cmp_u.st(cc, vxQ, sQ - 3 * dQ);
                    if (MIB::out1) put_u<ST, true>(o1_u, x, vxT, yu, sT - 3 * dT, o1, gl.u, nx, ny, P, A.oma, cm);
                    if (MIB::ghost2) put_u<Q, true>(o2_u, x, vxQ, yu, sQ - 3 * dQ, o2, gl.u, nx, ny, P, A.oma, cm);
                    else o2_u.st(o2, vxQ, sQ - 3 * dQ);
                }
                if (e_col) {
                    const Q deta = continuity_rhs<CS, Q>(rb.dUdx, rb.dVdy, rb.e_r1, A.cont, sb.cs, A.gamma, A.Fet, A.sinv);
                    Q o1 = 0, o2 = 0, cc = 0;
                    if (MIB::comp) cc = cmp_e.ld(vxQ, sQ - 3 * dQ);
                    combine<MB>(Q(0), qpe_3, E0_2, deta, A.k2, o1, o2, cc);
                    if (MIB::comp) cmp_e.st(cc, vxQ, sQ - 3 * dQ);
                    if (MIB::out1) put_e<ST, true>(o1_e, x, vxT, yu, sT - 3 * dT, o1, gl.e, nx, ny, P, cm);
                    if (MIB::ghost2) put_e<Q, true>(o2_e, x, vxQ, yu, sQ - 3 * dQ, o2, gl.e, nx, ny, P, cm);
                    else o2_e.st(o2, vxQ, sQ - 3 * dQ);
                }
            }
            const int yv = r - 4;
            if (v_col && yv >= t.ya && yv < t.yb && yv <= ny - 1) {
                const Q dv = -(toQ<Q>(rb.qhu) + toQ<Q>(rb.py)) + toQ<Q>(A.Fyt * sb.Fy);
                Q o1 = 0, o2 = 0, cc = 0;
                if (MIB::comp) cc = cmp_v.ld(vxQ, sQ - 4 * dQ);
                combine<MB>(Q(0), qpv_4, V0_2, dv, A.k2, o1, o2, cc);
                if (MIB::comp) cmp_v.st(cc, vxQ, sQ - 4 * dQ);
                if (MIB::out1) put_v<ST, true>(o1_v, x, vxT, yv, sT - 4 * dT, o1, gl.v, nx, P, A.oma, cm);
                if (MIB::ghost2) put_v<Q, true>(o2_v, x, vxQ, yv, sQ - 4 * dQ, o2, gl.v, nx, P, A.oma, cm);
                else o2_v.st(o2, vxQ, sQ - 4 * dQ);
            }
        }
        if constexpr (PRHIST) { wu_3 = wu_2; wu_2 = wu_1; we_3 = we_2; we_2 = we_1; wv_4 = wv_3; wv_3 = wv_2; }
        wu_1 = wu_0; wv_2 = wv_1; wv_1 = wv_0; we_1 = we_0;
        su_1 = su_0; sv_2 = sv_1; sv_1 = sv_0; se_1 = se_0;
        H_2 = H_1; H_1 = H_0; Hs_2 = Hs_1; Hs_1 = Hs_0;
        st_2 = st_1; st_1 = sa; ss_2 = ss_1; ss_1 = sas;
    }
}

template <class F, int MA, int MB, bool CS>
__global__ void __launch_bounds__(256) fused_seam_kernel(StageArgs<F> A) {
    fused_seam_body<F, MA, MB, CS>(A, wave_index());
}

// the seam strips of every member of an ensemble in one launch
template <class F, int MA, int MB, bool CS>
__global__ void __launch_bounds__(256) fused_seam_batch(const StageArgs<F> *__restrict__ args,
                                                        const int *__restrict__ start, int nmem) {
    int m, wl;
    if (!batch_locate(start, nmem, wave_index(), m, wl)) return;
    const StageArgs<F> A = args[m];
    fused_seam_body<F, MA, MB, CS>(A, wl);
}

template <class F, int MA, int MB, bool CS, bool PER = false, int ADV = ADV_AH>
__global__ void __launch_bounds__(256) fused_kernel(StageArgs<F> A) {
    fused_body<F, MA, MB, CS, PER, ADV>(A, wave_index());
}

template <class F, int MA, int MB, bool CS, bool PER = false, int ADV = ADV_AH>
__global__ void __launch_bounds__(256) fused_batch(const StageArgs<F> *__restrict__ args, const int *__restrict__ start,
                                                   int nmem) {
    int m, wl;
    if (!batch_locate(start, nmem, wave_index(), m, wl)) return;
    const StageArgs<F> A = args[m];
    fused_body<F, MA, MB, CS, PER, ADV>(A, wl);
}

// ------------------------------------------------------------------ semi-implicit continuity
//
// continuity! of the SSPRK schemes (continuity.jl:72-98 with UVfluxes! recomputed from the
// updated u1, v1 and the stage's η1, :83-85), followed by the η combination of the scheme
// (time_integration.jl:103,143-146,183-189) and the η ghost copies.  Emits row r-1.  The
// SSPRK stage variables are Tprog throughout.
template <class F, bool PER, int M>
__global__ void __launch_bounds__(256) continuity_kernel(StageArgs<F> A) {
    using MI = Mode<M>;
    using R = typename F::R;
    using Q = typename F::Q;
    using ST = typename F::ST;
    const WaveTile t = tile_of<HALO>(wave_index(), A.nstrips, A.nseg, A.L, A.ny);
    if (!t.live) return;
    const int x = t.x, P = A.P, nx = A.nx, ny = A.ny;
    const bool e_col = t.lane >= HALO && t.lane < WAVE - HALO && x >= 1 && x <= nx;
    const GhostLanes gl = ghost_lanes<PER>(x, nx);
    const Arena ar(A.ar);
    const Fld<Q> in_u{&ar, A.in_u}, in_v{&ar, A.in_v}, in_e{&ar, A.in_e};
    const Fld<ST> fH{&ar, A.st.H};
    const Fld<Q> pr_e{&ar, A.pr_e}, ac_e{&ar, A.ac_e}, o1_e{&ar, A.o1_e}, o2_e{&ar, A.o2_e};
    // the static of the continuity variant, a row ahead (row r-1 of the iteration that uses it)
    const Fld<ST> fcs{&ar, A.cont == CONT_RELAX ? A.st.eta_ref : A.st.Feta};
    const uint32_t vxQ = colb<Q>(x), vxT = colb<ST>(x);
    const uint32_t dQ = (uint32_t)P * sizeof(Q), dT = (uint32_t)P * sizeof(ST);
    const R half = R(0.5);

    R v_1 = 0, e_1 = 0, h_1 = 0, U_1 = 0, V_2 = 0;
    Q we_1 = 0;
    int r = t.ya - 1;
    const int rlast = t.yb;
    uint32_t sQ = rowb<Q>(r, P), sT = rowb<ST>(r, P);
    Q nu = in_u.ld(vxQ, sQ), nv = in_v.ld(vxQ, sQ), ne = in_e.ld(vxQ, sQ), npe = 0, nae = 0;
    R nH = fH.template ldr<R>(vxT, sT);
    R ncs = A.cont == CONT_ITSELF ? R(0) : fcs.template ldr<R>(vxT, sT - dT);
    // two rows per trip, and loads on every row at rows clamped to the segment (as in the
    // fused pairs); an odd count of rows gets one more at the end, which stores nothing
    const int rend = rlast + ((rlast - r + 1) & 1);
    auto row = [&](const int r, const uint32_t sQ, const uint32_t sT) __attribute__((always_inline)) {
        const Q we_0 = ne, pe_1 = npe, ae_1 = nae;
        const R u_0 = toR<R>(nu), v_0 = toR<R>(nv), e_0 = toR<R>(ne), H_0 = nH, cs_1 = ncs;
        if (A.cont != CONT_ITSELF) ncs = fcs.template ldr<R>(vxT, sT);
        {
            const int rn = min(r + 1, rlast);
            nu = in_u.ld(vxQ, rowb<Q>(rn, P)); nv = in_v.ld(vxQ, rowb<Q>(rn, P)); ne = in_e.ld(vxQ, rowb<Q>(rn, P));
            nH = fH.template ldr<R>(vxT, rowb<ST>(rn, P));
            const uint32_t se = rowb<Q>(min(max(r, t.ya), t.yb - 1), P);
            if (MI::need_pr) npe = pr_e.ld(vxQ, se);
            if (MI::need_ac) nae = ac_e.ld(vxQ, se);
        }
        // UVfluxes!  rhs.jl:307-324 from the updated velocities
        const R h_0 = e_0 + H_0;
        const R U_0 = u_0 * (half * (sh_xp(h_0) + h_0));
        const R V_1 = v_1 * (half * (h_0 + h_1));
        const R Uxm_1 = sh_xm(U_1);
        const int y = r - 1;
        if (e_col && y >= t.ya && y < t.yb) {
            const Q deta = continuity_rhs<true, Q>(U_1 - Uxm_1, V_1 - V_2, e_1, A.cont, cs_1, A.gamma, A.Fet, A.sinv);
            Q o1 = 0, o2 = 0;
            Q nocc = 0;
            combine<M>(we_1, pe_1, ae_1, deta, A.k, o1, o2, nocc);
            if (M == M_S4_3) put_e<Q, PER>(o1_e, x, vxQ, y, sQ - dQ, o2, gl.e, nx, ny, P, A.cm);   // only 2/3 η + 1/3 η1' lives on
            else {
                put_e<Q, PER>(o1_e, x, vxQ, y, sQ - dQ, o1, gl.e, nx, ny, P, A.cm);
                if (MI::out2) put_e<Q, PER>(o2_e, x, vxQ, y, sQ - dQ, o2, gl.e, nx, ny, P, A.cm);
            }
        }
        v_1 = v_0; e_1 = e_0; we_1 = we_0; h_1 = h_0; U_1 = U_0; V_2 = V_1;
    };
    for (; r <= rend; r += 2, sQ += 2 * dQ, sT += 2 * dT) {
        row(r, sQ, sT);
        row(r + 1, sQ + dQ, sT + dT);
    }
}

// ------------------------------------------------------------------ kept advection terms
//
// nstep_advcor > 0 (time_integration.jl:204-211, rhs.jl:36-38): q and KEu+KEv, the part of
// the Bernoulli potential the kinetic energy makes, are formed from a state (u, v, η) in
// format SI and kept for the stages of the next steps.  One thread per T point (x, y) of
// the η grid with its ghost ring (x = 0..nx+1, y = 0..ny+1) forms KEu(x,y) + KEv(x,y) as
// bernoulli! reads them (Ix!(KEu,u²), Iy!(KEv,v²), rhs.jl:107-109), and for x <= nx, y <= ny
// also q at the corner (x, y) (PV!, PV_advection.jl:2-18, with Ixy!(h_q,h)); the same
// operations in the same order as the stage kernels' right-hand side.  WQ: also q (the
// update of advection_coriolis!; bottom_drag_quadratic! refreshes KEu, KEv alone,
// bottom_drag.jl:25-38).  INIT: the thickness of the initial state, thickness!(h,η,H) on the
// Tprog η (time_integration.jl:21): η + H in the wider of T and Tprog, rounded to T.
template <class F, class SI, bool WQ, bool INIT>
__global__ void __launch_bounds__(256) advcor_kernel(StageArgs<F> A) {
    using R = typename F::R;
    using Q = typename F::Q;
    using ST = typename F::ST;
    const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x), y = (int)blockIdx.y;
    if (x > A.nx + 1 || y > A.ny + 1) return;
    const Arena ar(A.ar);
    const Fld<SI> u{&ar, A.in_u}, v{&ar, A.in_v}, e{&ar, A.in_e};
    const Fld<ST> H{&ar, A.st.H}, fq{&ar, A.st.fq}, qs{&ar, A.qs}, ks{&ar, A.ks};
    const R half = R(0.5), quarter = R(0.25);
    const int P = A.P;
    const uint32_t cm = colb<SI>(x - 1), c0 = colb<SI>(x), r0 = rowb<SI>(y, P), rm = rowb<SI>(y - 1, P);
    const R u0 = u.template ldr<R>(c0, r0), um = u.template ldr<R>(cm, r0);
    const R v0 = v.template ldr<R>(c0, r0), vm = v.template ldr<R>(c0, rm);
    const R uu0 = u0 * u0, uum = um * um;
    const R KEu = half * (uu0 + uum);
    const R KEv = half * (v0 * v0 + vm * vm);
    ks.st(KEu + KEv, colb<ST>(x), rowb<ST>(y, P));
    if (!WQ || x > A.nx || y > A.ny) return;
    const uint32_t c1 = colb<SI>(x + 1), r1 = rowb<SI>(y + 1, P);
    const uint32_t t0 = colb<ST>(x), t1 = colb<ST>(x + 1), q0 = rowb<ST>(y, P), q1 = rowb<ST>(y + 1, P);
    auto thick = [&](uint32_t ce, uint32_t re, uint32_t ch, uint32_t rh) -> R {
        if constexpr (INIT) return toR<R>(e.template ldr<Q>(ce, re) + toQ<Q>(H.template ldr<R>(ch, rh)));
        else return e.template ldr<R>(ce, re) + H.template ldr<R>(ch, rh);
    };
    const R h00 = thick(c0, r0, t0, q0), h10 = thick(c1, r0, t1, q0);
    const R h01 = thick(c0, r1, t0, q1), h11 = thick(c1, r1, t1, q1);
    const R hq = quarter * (h10 + h00) + quarter * (h11 + h01);
    const R dvdx = v.template ldr<R>(c1, r0) - v0;
    const R dudy = u.template ldr<R>(c0, r1) - u0;
    const R f = fq.template ldr<R>(scolb<ST>(x, A.st.fq_2d), srowb<ST>(y, P, A.st.fq_2d));
    qs.st((f + dvdx - dudy) / hq, t0, q0);
}

// q at the corner column x = nx for every corner row y = 0..ny (PV!, PV_advection.jl:15):
// the periodic AHβ!/AHγ! read q[end,·] for their first column (PV_advection.jl:110,126).
// SI is the storage of the stage input.
template <class F, class SI>
__device__ __forceinline__ void qend_body(const StageArgs<F> &A, int y) {
    using R = typename F::R;
    using ST = typename F::ST;
    if (y > A.ny) return;
    const Arena ar(A.ar);
    const Fld<SI> u{&ar, A.in_u}, v{&ar, A.in_v}, e{&ar, A.in_e};
    const Fld<ST> H{&ar, A.st.H}, fq{&ar, A.st.fq};
    const Fld<QES<F>> qe{&ar, A.qend};
    const R quarter = R(0.25);
    const int x = A.nx, P = A.P;
    const uint32_t c0 = colb<SI>(x), c1 = colb<SI>(x + 1), r0 = rowb<SI>(y, P), r1 = rowb<SI>(y + 1, P);
    const uint32_t t0 = colb<ST>(x), t1 = colb<ST>(x + 1), q0 = rowb<ST>(y, P), q1 = rowb<ST>(y + 1, P);
    const R h00 = e.template ldr<R>(c0, r0) + H.template ldr<R>(t0, q0);
    const R h10 = e.template ldr<R>(c1, r0) + H.template ldr<R>(t1, q0);
    const R h01 = e.template ldr<R>(c0, r1) + H.template ldr<R>(t0, q1);
    const R h11 = e.template ldr<R>(c1, r1) + H.template ldr<R>(t1, q1);
    const R hq = quarter * (h00 + h10) + quarter * (h01 + h11);
    const R dvdx = v.template ldr<R>(c1, r0) - v.template ldr<R>(c0, r0);
    const R dudy = u.template ldr<R>(c0, r1) - u.template ldr<R>(c0, r0);
    const R f = fq.template ldr<R>(scolb<ST>(x, A.st.fq_2d), srowb<ST>(y, P, A.st.fq_2d));
    qe.st((f + dvdx - dudy) / hq, 0, (uint32_t)(y + OY) * (uint32_t)sizeof(QES<F>));
}

template <class F, class SI>
__global__ void qend_kernel(StageArgs<F> A) {
    qend_body<F, SI>(A, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}

// all members: start[m] is member m's first block of 256 rows
template <class F, class SI>
__global__ void qend_batch(const StageArgs<F> *__restrict__ args, const int *__restrict__ start, int nmem) {
    int m, bl;
    if (!batch_locate(start, nmem, (int)blockIdx.x, m, bl)) return;
    const StageArgs<F> A = args[m];
    qend_body<F, SI>(A, bl * (int)blockDim.x + (int)threadIdx.x);
}

}  // namespace swm
