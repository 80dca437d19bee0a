// swm_device.h -- gfx950 kernels of the ShallowWaters.jl time step.
//
// Device layout.  Every field, whatever its staggering, lives in one padded 2-D layout
// indexed by T-cell coordinates (x, y): element (x, y) is at ptr0[x + y*P], ptr0 pointing
// at (0, 0).  A field's staggering is implicit:
//   T-point  (η, h, H, p, ...)  centre of cell (x, y)
//   U-point  (u, U, du, Fx ...) right face of cell (x, y)
//   V-point  (v, V, dv, Fy ...) top face of cell (x, y)
//   Q-point  (q, f_q, ...)      top-right corner of cell (x, y)
// so every reference stencil becomes a fixed offset pattern in {-1,0,1}^2 and the
// reference's ep-dependent index shifts disappear into the mapping Julia index <-> (x, y):
//   u[ii,jj] -> (ii-2-ep, jj-2)   v[ii,jj] -> (ii-2, jj-2)   η[ii,jj] -> (ii-1, jj-1)
//   q[ii,jj] -> (ii-1, jj-1)      sst[ii,jj] -> (ii-1, jj)
// The padding (x from -64, y from -4) keeps every lane's load in bounds without checks;
// padding is zero and never written.
//
// Kernel shape.  One wavefront owns a strip of 64 consecutive x (lane = x - xs) and a
// segment of rows, and marches upward in y keeping the few rows each stencil needs in
// registers.  Neighbours in x come from the adjacent lanes (cross-lane moves, no LDS, no
// barriers); every stencil of the step has an x-radius of 2, so lanes 2..61 of a strip
// produce output and strips advance by 60 columns.  Each output is written once, together
// with the ghost copies the reference's ghost_points! makes from it.
//
// Arithmetic follows the reference term by term in the same order (compiled with
// -ffp-contract=off; IEEE division and sqrt), so results are bitwise those of the CPU
// restatement for Float64 and Float32.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace swm {

constexpr int OX = 64;       // left padding in x (elements)
constexpr int OY = 4;        // bottom padding in y (rows)
constexpr int WAVE = 64;
constexpr int HALO = 2;      // x-radius of every fused stencil
constexpr int WOUT = WAVE - 2 * HALO;   // output columns per strip
constexpr int WAVES_PER_BLOCK = 4;

// ------------------------------------------------------------------ cross-lane moves
// value of lane+1 (x+1) and lane-1 (x-1); the edge lanes receive junk that only feeds
// the two halo lanes on each side.
__device__ __forceinline__ float sh_xp(float v) { return __shfl_down(v, 1); }
__device__ __forceinline__ float sh_xm(float v) { return __shfl_up(v, 1); }
__device__ __forceinline__ double sh_xp(double v) { return __shfl_down(v, 1); }
__device__ __forceinline__ double sh_xm(double v) { return __shfl_up(v, 1); }

template <class R> __device__ __forceinline__ R rsqrt_exact(R x) { return sqrt(x); }

// ------------------------------------------------------------------ parameters

template <class R> struct Static {
    const R *H;                       // T-point, 2-D
    const R *Fx; int Fx_sx, Fx_sy;    // U-point; sx=0 when x-invariant (1-D in y)
    const R *Fy; int Fy_sx, Fy_sy;    // V-point
    const R *fq; int fq_sx, fq_sy;    // Q-point
};

template <class R> struct StageArgs {
    const R *in_u, *in_v, *in_e;      // stage input (u1 of the reference; Prog at stage 1)
    const R *pr_u, *pr_v, *pr_e;      // Prog: base of c = a + x*b (caxb!)
    R *acc_u, *acc_v, *acc_e;         // u0 accumulator (axb!), read-modify-write
    R *out_u, *out_v, *out_e;         // next stage input (u1), null at the last stage
    const R *qend;                    // periodic: q at corner x=nx per row (AHβ/AHγ wrap)
    Static<R> st;
    R a, b, Fxt, Fyt, g, oma;
    int nx, ny, P, L, nstrips, nseg;
};

template <class R> struct DissArgs {
    const R *u0, *v0, *e0;            // state after the RK stages (ghosts applied)
    R *out_u, *out_v;                 // u0 + Δt_diff*(drag + diffusion), with ghosts
    const R *H;
    R cD, rD, cS, nuB, dtd, oma;
    int nx, ny, P, L, nstrips, nseg;
};

// ------------------------------------------------------------------ ghost copies on store
// ghost_points.jl:47-133: each interior value is written together with the halo cells
// the reference copies from it.

template <class R, bool PER>
__device__ __forceinline__ void store_u(R *b, int x, int y, R w, int nx, int ny, int P, R oma) {
    b[x + y * P] = w;
    if (PER) {                                   // ghost_points_u_periodic!  :60-73
        if (x == nx - 2) b[-2 + y * P] = w;
        if (x == nx - 1) b[-1 + y * P] = w;
        if (x == 0) b[nx + y * P] = w;
        if (x == 1) b[nx + 1 + y * P] = w;
    } else {                                     // ghost_points_u_nonperiodic!  :47-57
        R gw = oma * w;
        if (y == 1) b[x] = gw;
        if (y == 2) b[x - P] = gw;
        if (y == ny) b[x + (ny + 1) * P] = gw;
        if (y == ny - 1) b[x + (ny + 2) * P] = gw;
    }
}

template <class R, bool PER>
__device__ __forceinline__ void store_v(R *b, int x, int y, R w, int nx, int P, R oma) {
    b[x + y * P] = w;
    if (PER) {                                   // ghost_points_v_periodic!  :89-99
        if (x == nx - 1) b[-1 + y * P] = w;
        if (x == nx) b[y * P] = w;
        if (x == 1) b[nx + 1 + y * P] = w;
        if (x == 2) b[nx + 2 + y * P] = w;
    } else {                                     // ghost_points_v_nonperiodic!  :76-86
        R gw = oma * w;
        if (x == 1) b[y * P] = gw;
        if (x == 2) b[-1 + y * P] = gw;
        if (x == nx) b[nx + 1 + y * P] = gw;
        if (x == nx - 1) b[nx + 2 + y * P] = gw;
    }
}

template <class R, bool PER>
__device__ __forceinline__ void store_e(R *b, int x, int y, R w, int nx, int ny, int P) {
    // ghost_points_η_(non)periodic!  :102-133: x-copies first, then y-copies of whole rows
    b[x + y * P] = w;
    const bool gy0 = y == 1, gy1 = y == ny;
    if (gy0) b[x] = w;
    if (gy1) b[x + (ny + 1) * P] = w;
    if (x == (PER ? nx : 1)) {
        b[y * P] = w;
        if (gy0) b[0] = w;
        if (gy1) b[(ny + 1) * P] = w;
    }
    if (x == (PER ? 1 : nx)) {
        b[nx + 1 + y * P] = w;
        if (gy0) b[nx + 1] = w;
        if (gy1) b[nx + 1 + (ny + 1) * P] = w;
    }
}

// ------------------------------------------------------------------ wave -> (strip, segment)

struct WaveTile {
    int lane, x, ya, yb;
    bool live;
};

__device__ __forceinline__ WaveTile wave_tile(int nstrips, int nseg, int L, int ny) {
    WaveTile t;
    const int w = blockIdx.x * WAVES_PER_BLOCK + (int)(threadIdx.x / WAVE);
    t.lane = (int)(threadIdx.x % WAVE);
    t.live = w < nstrips * nseg;
    const int s = w % nstrips, g = w / nstrips;
    t.x = -HALO + WOUT * s + t.lane;
    t.ya = 1 + g * L;
    t.yb = min(t.ya + L, ny + 1);
    return t;
}

// ------------------------------------------------------------------ RK stage (K1)
//
// One stage of the RK loop, time_integration.jl:55-78:
//   rhs_nonlinear!  rhs.jl:24-51  (UVfluxes!, advection_coriolis!, PV_ArakawaHsu!,
//                                   bernoulli!, ∂x/∂y of p, momentum_u!/momentum_v!)
//   continuity!     continuity.jl:72-98 (continuity_itself!)
//   caxb!/axb!      time_integration.jl:69-77
//   ghost_points!   of the next stage input (and of u0 after the last stage, :198)
// At input row r the wave forms U(r), then the row-(r-1) quantities h_v, V, h_q, q, KE,
// p and the Arakawa-Hsu weights, and emits du, dη at row r-1 and dv at row r-2.
template <class R, bool PER, bool FIRST, bool LAST>
__global__ void __launch_bounds__(256) rk_stage_kernel(StageArgs<R> A) {
    const WaveTile t = wave_tile(A.nstrips, A.nseg, A.L, A.ny);
    if (!t.live) return;
    const int x = t.x, P = A.P, nx = A.nx, ny = A.ny;
    const bool lane_ok = t.lane >= HALO && t.lane < WAVE - HALO;
    const bool u_col = lane_ok && (PER ? (x >= 0 && x < nx) : (x >= 1 && x <= nx - 1));
    const bool v_col = lane_ok && x >= 1 && x <= nx;
    const bool e_col = v_col;
    const bool qfix = PER && x == 0;

    const R half = R(0.5), quarter = R(0.25), k12 = R(1.0 / 12.0);
    const R *fq_row = A.st.fq + x * A.st.fq_sx;
    const R *Fx_row = A.st.Fx + x * A.st.Fx_sx;
    const R *Fy_row = A.st.Fy + x * A.st.Fy_sx;

    // row history (suffix _1 = r-1, _2 = r-2)
    R u_1 = 0, v_1 = 0, v_2 = 0, e_1 = 0, h_1 = 0, hx_1 = 0;
    R U_1 = 0, U_2 = 0, Uxm_2 = 0, V_2 = 0, Vxp_2 = 0;
    R q_2 = 0, qxp_2 = 0, qxm_2 = 0, qaxm_2 = 0, qb_2 = 0, p_2 = 0, qend_2 = 0;

    // Loads run one row ahead of their use: the stencil inputs of row r+1 and the
    // pointwise operands (Prog, u0) of the rows emitted next iteration are in flight
    // while row r is computed.
    int r = t.ya - 1;
    const int rlast = t.yb + 1;
    const R *pu = A.in_u + x + r * P, *pv = A.in_v + x + r * P, *pe = A.in_e + x + r * P;
    const R *pH = A.st.H + x + r * P;
    R nu = *pu, nv = *pv, ne = *pe, nH = *pH;
    R npu = 0, nau = 0, npe = 0, nae = 0, npv = 0, nav = 0;   // operands of rows r-1 (u, η) and r-2 (v)
    for (; r <= rlast; ++r) {
        const R u_0 = nu, v_0 = nv, e_0 = ne, H_0 = nH;
        const R pu_1 = npu, au_1 = nau, pe_1 = npe, ae_1 = nae, pv_2 = npv, av_2 = nav;
        if (r < rlast) {
            pu += P; pv += P; pe += P; pH += P;
            nu = *pu; nv = *pv; ne = *pe; nH = *pH;
        }
        if (!FIRST && r >= t.ya && r < t.yb) {               // row r is emitted (u, η) next iteration
            if (!LAST) { npu = A.pr_u[x + r * P]; npe = A.pr_e[x + r * P]; }
            nau = A.acc_u[x + r * P];
            nae = A.acc_e[x + r * P];
        }
        if (!FIRST && r - 1 >= t.ya && r - 1 < t.yb) {       // row r-1 of v is emitted next iteration
            if (!LAST) npv = A.pr_v[x + (r - 1) * P];
            nav = A.acc_v[x + (r - 1) * P];
        }
        // thickness!  rhs.jl:131-139 ; Ix!(h_u)  interpolations.jl:3-14 ; Uflux!  rhs.jl:277-291
        const R h_0 = e_0 + H_0;
        const R hx_0 = sh_xp(h_0);
        const R U_0 = u_0 * (half * (hx_0 + h_0));
        // row r-1: Iy!(h_v), Vflux!  rhs.jl:294-304
        const R V_1 = v_1 * (half * (h_0 + h_1));
        // Ixy!(h_q,h)  interpolations.jl:33-45
        const R hq_1 = quarter * (h_1 + hx_1) + quarter * (h_0 + hx_0);
        // ∂x!(dvdx,v), ∂y!(dudy,u)  gradients.jl:3-25 ; PV!  PV_advection.jl:2-18
        const R dvdx_1 = sh_xp(v_1) - v_1;
        const R dudy_1 = u_0 - u_1;
        const R q_1 = (fq_row[(r - 1) * A.st.fq_sy] + dvdx_1 - dudy_1) / hq_1;
        // speed!, Ix!(KEu), Iy!(KEv)  rhs.jl:107-109 ; bernoulli!  rhs.jl:162-180
        const R uu_1 = u_1 * u_1;
        const R KEu_1 = half * (uu_1 + sh_xm(uu_1));
        const R KEv_1 = half * (v_1 * v_1 + v_2 * v_2);
        const R p_1 = half * (KEu_1 + KEv_1) + A.g * e_1;
        // AHα!..AHδ!  PV_advection.jl:88-146 at row r-1 (β, γ wrap to q[end,·] at x=0)
        const R qxp_1 = sh_xp(q_1);
        R qxm_1 = sh_xm(q_1);
        R qend_1 = 0;
        if (PER) {
            qend_1 = A.qend[r - 1];
            if (qfix) qxm_1 = qend_1;
        }
        const R qa_1 = k12 * (q_2 + q_1 + qxp_1);
        const R qb_1 = k12 * (q_2 + q_1 + qxm_1);
        // the wrapped first column of AHγ! sums (q[1,j] + q[1,j+1]) + q[end,j] (:126), the
        // other columns (q[i,j] + q[i-1,j]) + q[i,j+1] (:128)
        const R qg_1 = (PER && qfix) ? k12 * (q_2 + q_1 + qend_2) : k12 * (q_2 + qxm_2 + q_1);
        const R qd_1 = k12 * (q_2 + q_1 + qxp_2);

        // du at row r-1: PV_ArakawaHsu! qhv  PV_advection.jl:73 ; momentum_u!  rhs.jl:242
        const R Vxp_1 = sh_xp(V_1);
        const R p_xp1 = sh_xp(p_1);
        const int yu = r - 1;
        if (u_col && yu >= t.ya && yu < t.yb) {
            const R qhv = qa_1 * Vxp_1 + qb_1 * V_1 + qg_1 * V_2 + qd_1 * Vxp_2;
            const R du = (qhv - (p_xp1 - p_1)) + A.Fxt * Fx_row[yu * A.st.Fx_sy];
            const R base = FIRST ? u_1 : pu_1;
            const R acc = FIRST ? base : au_1;
            if (!LAST) store_u<R, PER>(A.out_u, x, yu, base + A.b * du, nx, ny, P, A.oma);
            const R a0 = acc + A.a * du;
            if (LAST) store_u<R, PER>(A.acc_u, x, yu, a0, nx, ny, P, A.oma);
            else A.acc_u[x + yu * P] = a0;
        }
        // dη at row r-1: ∂x!(dUdx,U), ∂y!(dVdy,V), continuity_itself!  continuity.jl:52-68
        const R Uxm_1 = sh_xm(U_1);
        if (e_col && yu >= t.ya && yu < t.yb) {
            const R deta = -((U_1 - Uxm_1) + (V_1 - V_2));
            const R base = FIRST ? e_1 : pe_1;
            const R acc = FIRST ? base : ae_1;
            if (!LAST) store_e<R, PER>(A.out_e, x, yu, base + A.b * deta, nx, ny, P);
            const R a0 = acc + A.a * deta;
            if (LAST) store_e<R, PER>(A.acc_e, x, yu, a0, nx, ny, P);
            else A.acc_e[x + yu * P] = a0;
        }
        // dv at row r-2: PV_ArakawaHsu! qhu  PV_advection.jl:81 ; momentum_v!  rhs.jl:271
        const R qdxm_1 = sh_xm(qd_1);
        const int yv = r - 2;
        if (v_col && yv >= t.ya && yv < t.yb && yv <= ny - 1) {
            const R qhu = qaxm_2 * Uxm_2 + qb_2 * U_2 + qg_1 * U_1 + qdxm_1 * Uxm_1;
            const R dv = -(qhu + (p_1 - p_2)) + A.Fyt * Fy_row[yv * A.st.Fy_sy];
            const R base = FIRST ? v_2 : pv_2;
            const R acc = FIRST ? base : av_2;
            if (!LAST) store_v<R, PER>(A.out_v, x, yv, base + A.b * dv, nx, P, A.oma);
            const R a0 = acc + A.a * dv;
            if (LAST) store_v<R, PER>(A.acc_v, x, yv, a0, nx, P, A.oma);
            else A.acc_v[x + yv * P] = a0;
        }
        // advance the row history
        const R qaxm_1 = sh_xm(qa_1);
        u_1 = u_0; v_2 = v_1; v_1 = v_0; e_1 = e_0; h_1 = h_0; hx_1 = hx_0;
        U_2 = U_1; U_1 = U_0; Uxm_2 = Uxm_1; V_2 = V_1; Vxp_2 = Vxp_1;
        q_2 = q_1; qxp_2 = qxp_1; qxm_2 = qxm_1; qaxm_2 = qaxm_1; qb_2 = qb_1; p_2 = p_1;
        qend_2 = qend_1;
    }
}

// q at the corner column x = nx for every corner row y = 0..ny (PV!, PV_advection.jl:15):
// the periodic AHβ!/AHγ! read q[end,·] for their first column (PV_advection.jl:110,126).
template <class R>
__global__ void qend_kernel(const R *u, const R *v, const R *e, const R *H, const R *fq, int fq_sx, int fq_sy,
                            R *qend, int nx, int ny, int P) {
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y > ny) return;
    const R half = R(0.5), quarter = R(0.25);
    (void)half;
    const int x = nx;
    const R h00 = e[x + y * P] + H[x + y * P], h10 = e[x + 1 + y * P] + H[x + 1 + y * P];
    const R h01 = e[x + (y + 1) * P] + H[x + (y + 1) * P], h11 = e[x + 1 + (y + 1) * P] + H[x + 1 + (y + 1) * P];
    const R hq = quarter * (h00 + h10) + quarter * (h01 + h11);
    const R dvdx = v[x + 1 + y * P] - v[x + y * P];
    const R dudy = u[x + (y + 1) * P] - u[x + y * P];
    qend[y] = (fq[x * fq_sx + y * fq_sy] + dvdx - dudy) / hq;
}

// ------------------------------------------------------------------ dissipation (K2)
//
// Once per step, time_integration.jl:216-221:
//   bottom_drag!   bottom_drag.jl:2-119 (quadratic, linear or none)
//   diffusion!     diffusion.jl:2-64 (Smagorinsky or constant biharmonic)
//   add_drag_diff_tendencies!  diffusion.jl:221-251
//   ghost_points_uv!           ghost_points.jl:202-216
// Emits u and v at row r-2 when input row r arrives.
template <class R, bool PER, int DRAG, bool SMAG>
__global__ void __launch_bounds__(256) dissipation_kernel(DissArgs<R> A) {
    const WaveTile t = wave_tile(A.nstrips, A.nseg, A.L, A.ny);
    if (!t.live) return;
    const int x = t.x, P = A.P, nx = A.nx, ny = A.ny;
    const bool lane_ok = t.lane >= HALO && t.lane < WAVE - HALO;
    const bool u_col = lane_ok && (PER ? (x >= 0 && x < nx) : (x >= 1 && x <= nx - 1));
    const bool v_col = lane_ok && x >= 1 && x <= nx;
    const R half = R(0.5), quarter = R(0.25);

    R u_1 = 0, u_2 = 0, uxm_1 = 0, v_1 = 0, v_2 = 0, h_1 = 0, h_2 = 0, sk_1 = 0, sk_2 = 0;
    R DT_1 = 0, DSq_2 = 0, DSqxm_2 = 0, nu_2 = 0, nuxp_2 = 0, Lu_2 = 0, Lv_2 = 0, Lvxp_2 = 0;
    R S11_2 = 0, S12_3 = 0, S22_2 = 0;

    int r = t.ya - 2;
    const int rlast = t.yb + 1;
    const R *pu = A.u0 + x + r * P, *pv = A.v0 + x + r * P, *pe = A.e0 + x + r * P, *pH = A.H + x + r * P;
    R nu = *pu, nv = *pv, ne = 0, nH = 0;              // row r+1 is loaded while row r is computed
    if (DRAG == 0) { ne = *pe; nH = *pH; }
    for (; r <= rlast; ++r) {
        const R u_0 = nu, v_0 = nv, e_0 = ne, H_0 = nH;
        if (r < rlast) {
            pu += P; pv += P;
            nu = *pu; nv = *pv;
            if (DRAG == 0) { pe += P; pH += P; ne = *pe; nH = *pH; }
        }
        const R uxm_0 = sh_xm(u_0);
        R h_0 = 0, sk_0 = 0;
        if (DRAG == 0) {
            // bottom_drag_quadratic!  bottom_drag.jl:20-76 (thickness, speed, KEu, KEv, sqrtKE)
            h_0 = e_0 + H_0;
            const R KEu_0 = half * (u_0 * u_0 + uxm_0 * uxm_0);
            const R KEv_0 = half * (v_0 * v_0 + v_1 * v_1);
            sk_0 = sqrt(KEu_0 + KEv_0);
        }
        // stress_tensor!  diffusion.jl:114-122: ∇²!  gradients.jl:29-42 at row r-1
        const R uxp_1 = sh_xp(u_1);
        const R Lu_1 = ((uxm_1 - u_1) + (uxp_1 - u_1)) + ((u_2 - u_1) + (u_0 - u_1));
        const R vxp_1 = sh_xp(v_1), vxm_1 = sh_xm(v_1);
        const R Lv_1 = ((vxm_1 - v_1) + (vxp_1 - v_1)) + ((v_2 - v_1) + (v_0 - v_1));
        R DT_0 = 0, nu_1, nuq_2, DSq_1 = 0, DSqxm_1 = 0;
        if (SMAG) {
            // smagorinsky_coeff!  diffusion.jl:68-111
            const R dudx_0 = u_0 - uxm_0, dvdy_0 = v_0 - v_1;
            const R s0 = dudx_0 + dvdy_0;
            DT_0 = s0 * s0;
            const R s1 = (u_0 - u_1) + (vxp_1 - v_1);
            DSq_1 = s1 * s1;
            DSqxm_1 = sh_xm(DSq_1);
            const R DS_1 = quarter * (DSqxm_2 + DSq_2) + quarter * (DSqxm_1 + DSq_1);
            nu_1 = A.cS * sqrt(DS_1 + DT_1);
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
        if (u_col && y >= t.ya && y < t.yb)
            store_u<R, PER>(A.out_u, x, y, u_2 + A.dtd * (Bu + LLu1 + LLu2), nx, ny, P, A.oma);
        if (v_col && y >= t.ya && y < t.yb && y <= ny - 1)
            store_v<R, PER>(A.out_v, x, y, v_2 + A.dtd * (Bv + LLv1 + LLv2), nx, P, A.oma);
        // advance
        u_2 = u_1; u_1 = u_0; uxm_1 = uxm_0; v_2 = v_1; v_1 = v_0; h_2 = h_1; h_1 = h_0; sk_2 = sk_1; sk_1 = sk_0;
        DT_1 = DT_0; DSq_2 = DSq_1; DSqxm_2 = DSqxm_1; nu_2 = nu_1; Lu_2 = Lu_1; Lv_2 = Lv_1; Lvxp_2 = Lvxp_1;
        S11_2 = S11_1; S12_3 = S12_2; S22_2 = S22_1;
    }
}

// ------------------------------------------------------------------ tracer (K3)
//
// tracer!  tracer_advection.jl:1-44 when i % nadvstep == 0: departure! (:52-85, with
// backtraj! :96-123 and interp_uv! :130-173 gathering the mid-point velocities um, vm
// interpolated to the T-grid on the fly), adv_sst! (:179-206), ghost_points_sst! and the
// copy back to sst.  One thread per T-cell.  idx(k,l) maps the reference's absolute
// indices of u_T/v_T/sst back to (x,y).
template <class R, bool PER>
__global__ void tracer_kernel(const R *u, const R *v, const R *um, const R *vm, const R *sst, R *ssti,
                              R hdu, R hdv, R du_, R dv_, int nx, int ny, int P, int ep, int consume,
                              R jSST, R SSTmin, int *bad) {
    const int x = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = 1 + (int)blockIdx.y;
    if (x > nx) return;
    const R half = R(0.5), one = R(1.0);
    // Ix!(u_T,u): u_T(x,y) = ½(u(x) + u(x-1)); Iy!(v_T,v): v_T(x,y) = ½(v(y) + v(y-1))
    const R uT = half * (u[x + y * P] + u[x - 1 + y * P]);
    const R vT = half * (v[x + y * P] + v[x + (y - 1) * P]);
    const R xd0 = -hdu * uT, yd0 = -hdv * vT;
    // interp_uv! on um_T (x-index k <-> T-column k-1-ep, y-index l <-> row l-2) and
    // vm_T (k <-> column k-2, l <-> row l-1)
    auto interp = [&](const R *f, bool is_u, R xx, R yy) -> R {
        const int muv = is_u ? nx + 2 + ep : nx + 4, nuv = is_u ? ny + 4 : ny + 2;
        const int xh1 = is_u ? 1 + ep : 2, xh2 = is_u ? 1 : 2, yh1 = is_u ? 2 : 1;
        const R fx = floor(xx), fy = floor(yy);
        if (!isfinite(fx) || !isfinite(fy)) { atomicOr(bad, 1); return R(0); }
        const long xi = (long)fx, yi = (long)fy;
        long k = xi + x + xh1, l = yi + y + yh1;
        R x0 = xx - R((double)xi), y0 = yy - R((double)yi);
        if (PER) {
            if (k < 1 + xh1) k = k + muv - xh1 - xh2;
            else if (k > muv - 1) k = k - muv + xh1 + xh2;
        } else {
            if (k < 1 + xh1) { k = 1 + xh1; x0 = R(0); }
            else if (k > muv - 1) { k = muv - 1; x0 = one; }
        }
        if (l < 1 + yh1) { l = 1 + yh1; y0 = R(0); }
        else if (l > nuv - 1) { l = nuv - 1; y0 = one; }
        if (k < 1 || k + 1 > muv) { atomicOr(bad, 2); return R(0); }
        // value of f_T at reference index (kk, ll)
        auto fT = [&](long kk, long ll) -> R {
            if (is_u) { const int cx = (int)kk - 1 - ep, cy = (int)ll - 2; return half * (f[cx + cy * P] + f[cx - 1 + cy * P]); }
            const int cx = (int)kk - 2, cy = (int)ll - 1;
            return half * (f[cx + cy * P] + f[cx + (cy - 1) * P]);
        };
        const R f00 = fT(k, l), f10 = fT(k + 1, l), f01 = fT(k, l + 1), f11 = fT(k + 1, l + 1);
        return f00 * (one - x0) * (one - y0) + f10 * x0 * (one - y0) + f01 * (one - x0) * y0 + f11 * x0 * y0;
    };
    const R ui = interp(um, true, xd0, yd0);
    const R vi = interp(vm, false, xd0, yd0);
    const R xd = -du_ * ui, yd = -dv_ * vi;
    // adv_sst!  (sst index k <-> column k-1, l <-> row l; m = nx+2, n = ny, halo (1,0))
    const R fx = floor(xd), fy = floor(yd);
    if (!isfinite(fx) || !isfinite(fy)) { atomicOr(bad, 1); return; }
    const long xi = (long)fx, yi = (long)fy;
    const int m = nx + 2, n = ny;
    long k = xi + x + 1, l = yi + y;
    R x0 = xd - R((double)xi), y0 = yd - R((double)yi);
    if (PER) {
        if (k < 2) k = k + m - 2;
        else if (k > m - 1) k = k - m + 2;
    } else {
        if (k < 2) { k = 2; x0 = R(0); }
        else if (k > m - 1) { k = m - 1; x0 = one; }
    }
    if (l < 1) { l = 1; y0 = R(0); }
    else if (l > n - 1) { l = n - 1; y0 = one; }
    if (k < 1 || k + 1 > m) { atomicOr(bad, 2); return; }
    auto S = [&](long kk, long ll) -> R { return sst[(int)kk - 1 + (int)ll * P]; };
    R s = S(k, l) * (one - x0) * (one - y0) + S(k + 1, l) * x0 * (one - y0) + S(k, l + 1) * (one - x0) * y0 +
          S(k + 1, l + 1) * x0 * y0;
    if (consume) s += jSST * (SSTmin - s);       // tracer_consumption!  :262-275
    // ghost_points_sst!  ghost_points.jl:136-179 (halosstx=1, halossty=0)
    ssti[x + y * P] = s;
    if (x == (PER ? nx : 1)) ssti[y * P] = s;
    if (x == (PER ? 1 : nx)) ssti[nx + 1 + y * P] = s;
}

// ------------------------------------------------------------------ small kernels

// ghost_points! of the uploaded state (ghost_points.jl:182-242), in the reference order:
// phase 0 writes the x-copies, phase 1 the y-copies (which read the x-copies for η).
template <class R, bool PER>
__global__ void ghost_kernel(R *u, R *v, R *e, R *sst, int nx, int ny, int P, R oma, int phase) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (phase == 0) {
        // rows of u (periodic: x-copies), v (x-copies), η (x-copies), sst (x-copies)
        const int y = k - 1;                     // y = -1 .. ny+2
        if (y < -1 || y > ny + 2) return;
        if (PER) {
            u[-2 + y * P] = u[nx - 2 + y * P];
            u[-1 + y * P] = u[nx - 1 + y * P];
            u[nx + y * P] = u[y * P];
            u[nx + 1 + y * P] = u[1 + y * P];
            if (y <= ny + 1) {
                v[-1 + y * P] = v[nx - 1 + y * P];
                v[y * P] = v[nx + y * P];
                v[nx + 1 + y * P] = v[1 + y * P];
                v[nx + 2 + y * P] = v[2 + y * P];
            }
            if (y >= 0 && y <= ny + 1) {
                e[y * P] = e[nx + y * P];
                e[nx + 1 + y * P] = e[1 + y * P];
            }
            if (y >= 1 && y <= ny) {
                sst[y * P] = sst[nx + y * P];
                sst[nx + 1 + y * P] = sst[1 + y * P];
            }
        } else {
            if (y <= ny + 1) {
                v[y * P] = oma * v[1 + y * P];
                v[-1 + y * P] = oma * v[2 + y * P];
                v[nx + 1 + y * P] = oma * v[nx + y * P];
                v[nx + 2 + y * P] = oma * v[nx - 1 + y * P];
            }
            if (y >= 0 && y <= ny + 1) {
                e[y * P] = e[1 + y * P];
                e[nx + 1 + y * P] = e[nx + y * P];
            }
            if (y >= 1 && y <= ny) {
                sst[y * P] = sst[1 + y * P];
                sst[nx + 1 + y * P] = sst[nx + y * P];
            }
        }
    } else {
        const int x = k - 2;                     // x = -2 .. nx+2
        if (x < -2 || x > nx + 2) return;
        const int xu_lo = PER ? -2 : -1, xu_hi = PER ? nx + 1 : nx + 1;
        if (!PER && x >= xu_lo && x <= xu_hi) {
            u[x] = oma * u[x + P];
            u[x - P] = oma * u[x + 2 * P];
            u[x + (ny + 1) * P] = oma * u[x + ny * P];
            u[x + (ny + 2) * P] = oma * u[x + (ny - 1) * P];
        }
        if (x >= 0 && x <= nx + 1) {
            e[x] = e[x + P];
            e[x + (ny + 1) * P] = e[x + ny * P];
        }
    }
}

// copy of a padded field (tracer mid-point snapshot um ← u0, tracer_advection.jl:12-17)
template <class R>
__global__ void copy_kernel(R *__restrict__ dst, const R *__restrict__ src, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

// any NaN in the listed fields (nan_detection!, feedback.jl:61-72); f[k] points at (0,0)
// of field k with rows y0[k]..y1[k] and columns x0[k]..x1[k]
struct NanArgs {
    const void *f[4];
    int x0[4], x1[4], y0[4], y1[4];
    int P, nf;
};
template <class R>
__global__ void nan_kernel(NanArgs A, int *flag) {
    const int y = (int)blockIdx.y + A.y0[blockIdx.z];
    const int k = blockIdx.z;
    if (y > A.y1[k]) return;
    const R *f = (const R *)A.f[k];
    bool bad = false;
    for (int x = A.x0[k] + (int)threadIdx.x; x <= A.x1[k]; x += blockDim.x) bad |= isnan(f[x + y * A.P]);
    if (__any(bad) && (threadIdx.x % WAVE) == 0) atomicOr(flag, 1);
}

// FP64 mass and energy over the interior (swm_reduce): block partial sums in double.
template <class R>
__global__ void reduce_kernel(const R *u, const R *v, const R *e, const R *H, int nx, int ny, int P, int ep, double g,
                              double *part) {
    __shared__ double sm[2][256];
    double mass = 0.0, en = 0.0;
    const int y = 1 + blockIdx.x;
    for (int x = 1 + (int)threadIdx.x; x <= nx; x += blockDim.x) {
        const double ee = (double)e[x + y * P];
        const double h = ee + (double)H[x + y * P];
        mass += ee;
        en += 0.5 * g * ee * ee;
        const double uu = (double)u[x + y * P], vv = (double)v[x + y * P];
        if (ep || x <= nx - 1) en += 0.5 * h * uu * uu;
        if (y <= ny - 1) en += 0.5 * h * vv * vv;
    }
    sm[0][threadIdx.x] = mass;
    sm[1][threadIdx.x] = en;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sm[0][threadIdx.x] += sm[0][threadIdx.x + s];
            sm[1][threadIdx.x] += sm[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sm[0][0];
        part[2 * blockIdx.x + 1] = sm[1][0];
    }
}

}  // namespace swm
