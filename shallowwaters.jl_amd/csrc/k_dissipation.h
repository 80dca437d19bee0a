// k_dissipation.h -- bottom drag + biharmonic diffusion, once per step (gfx950).
#pragma once

#include "swm_common.h"

namespace swm {

template <class F> struct DissArgs {
    using R = typename F::R;
    using Q = typename F::Q;
    ArenaRef ar;                      // the context's fields; the members below are origins in it
    uint32_t u0, v0, e0;              // state after the RK stages (ghosts applied), Tprog
    uint32_t out_u, out_v;            // u0 + Δt_diff*(drag + diffusion), with ghosts, Tprog
    uint32_t H;                       // T
    R cD, rD, cS, nuB;                // T
    R sinv;                           // 1/s for velocities carried scaled by s (1 unscaled)
    Q dtd, oma;                       // Δt_diff, one_minus_α (Tprog)
    int cm;                           // Tcomm
    int nx, ny, P, L, nstrips, nseg;
};

// time_integration.jl:201-221 with u0rhs = T(u0):
//   bottom_drag!   bottom_drag.jl:2-119 (quadratic, linear or none)
//   diffusion!     diffusion.jl:2-64 (Smagorinsky or constant biharmonic)
//   add_drag_diff_tendencies!  diffusion.jl:221-251 (in Tprog)
//   ghost_points_uv!           ghost_points.jl:202-216
// At input row r the wave forms the row r quantities (KE, sqrtKE, DT), the Laplacians and
// DS of row r-1, ν and the viscous tensor of rows r-1/r-2, and emits u and v at row r-2.
template <class F, bool PER, int DRAG, bool SMAG>
__device__ __forceinline__ void dissipation_body(const DissArgs<F> &A, int w) {
    using R = typename F::R;
    using Q = typename F::Q;
    using ST = typename F::ST;
    const WaveTile t = tile_of<HALO>(w, A.nstrips, A.nseg, A.L, A.ny);
    if (!t.live) return;
    const int x = t.x, P = A.P, nx = A.nx, ny = A.ny;
    const bool lane_ok = t.lane >= HALO && t.lane < WAVE - HALO;
    const bool u_col = lane_ok && (PER ? (x >= 0 && x < nx) : (x >= 1 && x <= nx - 1));
    const bool v_col = lane_ok && x >= 1 && x <= nx;
    const GhostLanes gl = ghost_lanes<PER>(x, nx);
    const Arena ar(A.ar);
    const Fld<Q> fu{&ar, A.u0}, fv{&ar, A.v0}, fe{&ar, A.e0}, ou{&ar, A.out_u}, ov{&ar, A.out_v};
    const Fld<ST> fH{&ar, A.H};
    const uint32_t vxQ = colb<Q>(x), vxT = colb<ST>(x);
    const uint32_t dQ = (uint32_t)P * sizeof(Q);
    const R half = R(0.5), quarter = R(0.25);

    R u_1 = 0, u_2 = 0, uxm_1 = 0, v_1 = 0, v_2 = 0, h_1 = 0, h_2 = 0, sk_1 = 0, sk_2 = 0;
    R DT_1 = 0, DSq_2 = 0, DSqxm_2 = 0, nu_2 = 0, nuxp_2 = 0, Lu_2 = 0, Lv_2 = 0, Lvxp_2 = 0;
    R S11_2 = 0, S12_3 = 0, S22_2 = 0;
    Q wu_1 = 0, wu_2 = 0, wv_1 = 0, wv_2 = 0;         // u0, v0 in Tprog for the Euler add

    int r = t.ya - 2;
    const int rlast = t.yb + 1;
    uint32_t sQ = rowb<Q>(r, P), sT = rowb<ST>(r, P);
    Q nu = fu.ld(vxQ, sQ), nv = fv.ld(vxQ, sQ), ne = 0;   // row r+1 is loaded while row r is computed
    R nH = 0;
    if (DRAG == 0) { ne = fe.ld(vxQ, sQ); nH = fH.template ldr<R>(vxT, sT); }
    // Two rows per trip (the row history stays in the same registers, no copies); an odd
    // count of rows gets one more at the end, which stores nothing (row yb).
    const int rend = rlast + ((rlast - r + 1) & 1);
    auto row = [&](const int r, const uint32_t sQ) __attribute__((always_inline)) {
        const Q wu_0 = nu, wv_0 = nv;
        const R u_0 = toR<R>(nu), v_0 = toR<R>(nv), e_0 = toR<R>(ne), H_0 = nH;
        {   // row r+1, clamped to the last row the march reads: a load on every row
            const int rn = min(r + 1, rlast);
            nu = fu.ld(vxQ, rowb<Q>(rn, P)); nv = fv.ld(vxQ, rowb<Q>(rn, P));
            if (DRAG == 0) { ne = fe.ld(vxQ, rowb<Q>(rn, P)); nH = fH.template ldr<R>(vxT, rowb<ST>(rn, P)); }
        }
        const R uxm_0 = sh_xm(u_0);
        R h_0 = 0, sk_0 = 0;
        if (DRAG == 0) {
            // bottom_drag_quadratic!  bottom_drag.jl:20-76 (thickness, speed, KEu, KEv, sqrtKE)
            h_0 = e_0 + H_0;
            const R KEu_0 = half * (u_0 * u_0 + uxm_0 * uxm_0);
            const R KEv_0 = half * (v_0 * v_0 + v_1 * v_1);
            sk_0 = nsqrt(KEu_0 + KEv_0) * A.sinv;
        }
        // stress_tensor!  diffusion.jl:114-122: ∇²!  gradients.jl:29-42 at row r-1
        const R uxp_1 = sh_xp(u_1);
        const R Lu_1 = ((uxm_1 - u_1) + (uxp_1 - u_1)) + ((u_2 - u_1) + (u_0 - u_1));
        const R vxp_1 = sh_xp(v_1), vxm_1 = sh_xm(v_1);
        const R Lv_1 = ((vxm_1 - v_1) + (vxp_1 - v_1)) + ((v_2 - v_1) + (v_0 - v_1));
        R DT_0 = 0, nu_1, nuq_2, DSq_1 = 0, DSqxm_1 = 0;
        if (SMAG) {
            // smagorinsky_coeff!  diffusion.jl:68-111 (tension uses +, :95)
            const R dudx_0 = u_0 - uxm_0, dvdy_0 = v_0 - v_1;
            const R s0 = dudx_0 + dvdy_0;
            DT_0 = s0 * s0;
            const R s1 = (u_0 - u_1) + (vxp_1 - v_1);
            DSq_1 = s1 * s1;
            DSqxm_1 = sh_xm(DSq_1);
            const R DS_1 = quarter * (DSqxm_2 + DSq_2) + quarter * (DSqxm_1 + DSq_1);
            nu_1 = A.cS * (nsqrt(DS_1 + DT_1) * A.sinv);
            const R nuxp_1 = sh_xp(nu_1);
            nuq_2 = quarter * (nu_2 + nuxp_2) + quarter * (nu_1 + nuxp_1);
            nuxp_2 = nuxp_1;
        } else {
            nu_1 = A.nuB;
            nuq_2 = A.nuB;
        }
        // viscous_tensor_*!  diffusion.jl:126-218
        const R S11_1 = nu_1 * (Lu_1 - sh_xm(Lu_1));
        const R S12_2 = nuq_2 * (Lu_1 - Lu_2);
        const R Lvxp_1 = sh_xp(Lv_1);
        const R S21_2 = nuq_2 * (Lvxp_2 - Lv_2);
        const R S22_1 = nu_1 * (Lv_1 - Lv_2);
        // ∂x!(LLu1,S11), ∂y!(LLu2,S12), ∂x!(LLv1,S21), ∂y!(LLv2,S22)  diffusion.jl:60-63 at row r-2
        const R LLu1 = sh_xp(S11_2) - S11_2;
        const R LLu2 = S12_2 - S12_3;
        const R LLv1 = S21_2 - sh_xm(S21_2);
        const R LLv2 = S22_1 - S22_2;
        R Bu = 0, Bv = 0;
        if (DRAG == 0) {
            const R sku = half * (sh_xp(sk_2) + sk_2), hu = half * (sh_xp(h_2) + h_2);
            Bu = A.cD * sku * u_2 / hu;
            const R skv = half * (sk_1 + sk_2), hv = half * (h_1 + h_2);
            Bv = A.cD * skv * v_2 / hv;
        } else if (DRAG == 1) {
            Bu = A.rD * u_2;
            Bv = A.rD * v_2;
        }
        const int y = r - 2;
        if (y >= t.ya && y < t.yb) {
            // u0 += Δt_diff*(Tprog(Bu) + Tprog(LLu1) + Tprog(LLu2))  diffusion.jl:237, :248 for v
            if (u_col) {
                const Q nu0 = wu_2 + A.dtd * (toQ<Q>(Bu) + toQ<Q>(LLu1) + toQ<Q>(LLu2));
                put_u<Q, PER>(ou, x, vxQ, y, sQ - 2 * dQ, nu0, gl.u, nx, ny, P, A.oma, A.cm);
            }
            if (v_col && y <= ny - 1) {
                const Q nv0 = wv_2 + A.dtd * (toQ<Q>(Bv) + toQ<Q>(LLv1) + toQ<Q>(LLv2));
                put_v<Q, PER>(ov, x, vxQ, y, sQ - 2 * dQ, nv0, gl.v, nx, P, A.oma, A.cm);
            }
        }
        // advance
        u_2 = u_1; u_1 = u_0; uxm_1 = uxm_0; v_2 = v_1; v_1 = v_0; h_2 = h_1; h_1 = h_0; sk_2 = sk_1; sk_1 = sk_0;
        wu_2 = wu_1; wu_1 = wu_0; wv_2 = wv_1; wv_1 = wv_0;
        DT_1 = DT_0; DSq_2 = DSq_1; DSqxm_2 = DSqxm_1; nu_2 = nu_1; Lu_2 = Lu_1; Lv_2 = Lv_1; Lvxp_2 = Lvxp_1;
        S11_2 = S11_1; S12_3 = S12_2; S22_2 = S22_1;
    };
    for (; r <= rend; r += 2, sQ += 2 * dQ) {
        row(r, sQ);
        row(r + 1, sQ + dQ);
    }
}

template <class F, bool PER, int DRAG, bool SMAG>
__global__ void __launch_bounds__(256) dissipation_kernel(DissArgs<F> A) {
    dissipation_body<F, PER, DRAG, SMAG>(A, wave_index());
}

template <class F, bool PER, int DRAG, bool SMAG>
__global__ void __launch_bounds__(256) dissipation_batch(const DissArgs<F> *__restrict__ args,
                                                         const int *__restrict__ start, int nmem) {
    int m, wl;
    if (!batch_locate(start, nmem, wave_index(), m, wl)) return;
    const DissArgs<F> A = args[m];
    dissipation_body<F, PER, DRAG, SMAG>(A, wl);
}

}  // namespace swm
