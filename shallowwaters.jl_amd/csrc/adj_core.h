// adj_core.h -- reverse-mode (adjoint) kernels of one Runge-Kutta time step (any of the
// reference's schemes: classic RK, SSPRK2, SSPRK3, 4SSPRK3).
//
// The time step x_{n+1} = F(x_n) of time_integration.jl:45-243 (T = Tprog) is
// differentiated by hand, operator by operator, in gather form: every adjoint value is a
// sum over the forward quantities that read the corresponding input, so each output cell
// is written by one lane and no atomics are needed; results do not depend on the launch
// shape.  Three kinds of kernels:
//
//   stage_adjoint   one RK stage: the transposed Jacobian of rhs! + continuity!
//                   (rhs.jl:2-51, continuity.jl:52-98) at the stage input s_k, applied to
//                   the stage's tendency adjoint, together with the stage combination's
//                   bookkeeping (caxb!/axb!, time_integration.jl:66-76).
//   diss_adjoint    bottom drag + biharmonic diffusion + add_drag_diff_tendencies!
//                   (time_integration.jl:216-221, bottom_drag.jl, diffusion.jl).
//   tracer_*        the semi-Lagrangian tracer (tracer_advection.jl:1-44): departure points,
//                   bilinear interpolation and their derivatives.
//
// Conventions at the kinks of the forward map: sqrt'(0) = 0 (Smagorinsky ν and drag √KE at
// rest); floor and the interpolation indices are constants; a clipped departure coordinate
// is a constant.  Ghost copies ("halo := c · interior", ghost_points.jl) are differentiated
// exactly: an adjoint field is carried with its halo cells, and a reader folds the halo
// into the interior cells they copy ("fold": λ(interior) += c · λ(halo)).
//
// The marches are written once against a wave policy W: on the GPU, W is one wavefront
// (64 lanes, one column each, neighbours in x by DPP lane shifts); a host build runs the
// same code on an emulated wavefront, lane by lane, so that the kernels can be checked
// without a GPU.  W provides the per-lane value type V, predicate M and integer I,
// xp / xm (the value of lane+1 / lane-1), sel, and row loads / masked stores in the padded
// field layout (element (x, y) at byte origin + ((x + 64) + (y + 4)·P)·sizeof(R)).
#pragma once

#include <stdint.h>

#include <type_traits>

#if defined(__HIP__)
#define SWM_HD __host__ __device__ __forceinline__
#else
#define SWM_HD inline
#endif

namespace swm {
namespace adj {

constexpr int AOX = 64;     // left padding (columns) of the field layout
constexpr int AOY = 4;      // bottom padding (rows)
constexpr int AWAVE = 64;   // lanes of a wave
constexpr int AX0 = -2;     // first column the adjoint marches produce (the u halo of the periodic case)

// x-halo lanes of each march: the x-radius of its dependency cone
constexpr int RT_HALO = 2;  // stage_adjoint
constexpr int DT_HALO = 3;  // diss_adjoint
// output rows of both marches: [-1, ny+3) (u rows -1 .. ny+2 of the nonperiodic halo)
constexpr int ADJ_Y0 = -1;

// the right-hand side forms of the stage adjoint (template argument SCH): the nonlinear
// dynamics with Arakawa-Hsu or Sadourny PV fluxes, the linear dynamics; ADJ_KEEP (with a
// nonlinear form): q and KEu+KEv kept from an earlier state (nstep_advcor > 0); ADJ_CONT
// (alone): the semi-implicit continuity of the SSP schemes only, the mass fluxes
// recomputed from the state (UVfluxes!, continuity.jl:83-85) whatever the dynamics
enum { ADJ_AH = 0, ADJ_SAD = 1, ADJ_LIN = 2, ADJ_KEEP = 4, ADJ_CONT = 8 };
SWM_HD int adj_y1(int ny) { return ny + 3; }

template <class T> SWM_HD T amin(T a, T b) { return a < b ? a : b; }
template <class T> SWM_HD T amax(T a, T b) { return a > b ? a : b; }

// ------------------------------------------------------------------ tiles

// wave w of a march: strips of AWAVE lanes with H halo lanes on each side, covering columns
// AX0 .. (own lanes H..AWAVE-H-1), times segments of L output rows covering [y0, y1)
struct Tile {
    int x_lane0;      // column of lane 0
    int ya, yb;       // output rows [ya, yb)
    bool live;
};
template <int H> SWM_HD Tile tile_of(int w, int nstrips, int nseg, int L, int y0, int y1) {
    Tile t;
    t.live = w >= 0 && w < nstrips * nseg;
    const int s = w % nstrips, g = w / nstrips;
    t.x_lane0 = AX0 - H + (AWAVE - 2 * H) * s;
    t.ya = y0 + g * L;
    t.yb = amin(t.ya + L, y1);
    return t;
}
// strips covering columns AX0 .. xmax
template <int H> SWM_HD int nstrips_for(int xmax) { return (xmax - AX0 + 1 + (AWAVE - 2 * H) - 1) / (AWAVE - 2 * H); }

// ------------------------------------------------------------------ arguments

// Fields are named by their byte origin in one arena (the context's device allocation).
struct UVE {
    uint32_t u, v, e;
};

template <class R> struct StageAdjArgs {
    UVE s;                 // stage input s_k (u1, v1, η1 after ghost_points!)
    uint32_t H, fq;        // static fields (f_q may be one value per row: fq_2d = 0)
    int fq_2d;
    uint32_t fu, fv;       // linear dynamics: f on the u- and v-grid (one value per row when fu_2d / fv_2d = 0)
    int fu_2d, fv_2d;
    uint32_t qk;           // kept q the stage read (nstep_advcor > 0), corners
    uint32_t lqk, lkk;     // λ of the kept q (corners) and KEu+KEv (T points), accumulated
    uint32_t qend;         // periodic: q at column nx for every corner row of s_k (1-D)
    UVE l0;                // λ of u0 after the stages (halo unfolded)
    UVE mu_in;             // μ_{k+1} = (∂R/∂s_{k+1})ᵀ λd_{k+1} (halo unfolded)
    UVE mu_x[2];           // the first stage: μ of the stages after its successor, oldest first (NX of them)
    UVE mu_out, a_out;     // μ_k; the first stage: λ of the step's input, A_1 + μ_1
    uint32_t lqend;        // periodic: λ of q(nx, y), 1-D
    uint32_t lfx;          // fxg: λ of the zonal wind stress Fx at the u points, accumulated
    R a, b;                // RKaΔt[k], RKbΔt[k] (b of the stage whose input is s_{k+1})
    R g, gamma, k12, oma;
    R fxt;                 // the step's factor of Fx in du (Ftime, forcing.jl:215-217; 1 unless seasonal)
    int cont, nx, ny, P, L, nstrips, nseg;
    int fxg;               // accumulate λFx += fxt·λdu (du = ... + Fxt·Fx, rhs.jl:242)
};

template <class R> struct DissAdjArgs {
    UVE s;                 // u0, v0, η0 before the dissipation (after ghost_points!)
    uint32_t H;
    UVE lin;               // λ after the dissipation (u, v halo unfolded; η passes through)
    UVE out;               // λ before the dissipation (halo unfolded)
    R cD, rD, cS, nuB, dtd, oma;
    int nx, ny, P, L, nstrips, nseg;
};

// ------------------------------------------------------------------ folds of ghost copies
//
// ghost_points.jl:47-133.  u (nonperiodic): rows 0, -1, ny+1, ny+2 copy (1-α)·u of rows 1,
// 2, ny, ny-1.  u (periodic): columns -2, -1, nx, nx+1 copy columns nx-2, nx-1, 0, 1.  v:
// columns 0, -1, nx+1, nx+2 copy (1-α)·v of columns 1, 2, nx, nx-1 (nonperiodic) or columns
// nx, nx-1, 1, 2 (periodic).  η: the halo ring copies its nearest interior cell (periodic:
// wrapped in x).  Tcomm = Tprog: the periodic copies are exact.  `l` is the cell's own value
// (row y of the field at o), loaded by the caller ahead of the fold.

template <class W, bool PER>
SWM_HD typename W::V fold_u(const W &w, uint32_t o, int y, typename W::V l, typename W::R oma, int nx, int ny) {
    if (PER) {
        const auto mlo = (w.x >= nx - 2) & (w.x <= nx - 1);
        const auto mhi = (w.x >= 0) & (w.x <= 1);
        l = W::sel(mlo, l + w.ld_at(o, w.x - nx, y, mlo), l);
        l = W::sel(mhi, l + w.ld_at(o, w.x + nx, y, mhi), l);
    } else {
        if (y == 1) l = l + oma * w.ld(o, 0);
        if (y == 2) l = l + oma * w.ld(o, -1);
        if (y == ny) l = l + oma * w.ld(o, ny + 1);
        if (y == ny - 1) l = l + oma * w.ld(o, ny + 2);
    }
    return l;
}

template <class W, bool PER>
SWM_HD typename W::V fold_v(const W &w, uint32_t o, int y, typename W::V l, typename W::R oma, int nx) {
    const auto m1 = w.x == 1, m2 = w.x == 2, m3 = w.x == nx, m4 = w.x == nx - 1;
    if (PER) {
        const auto mlo = m3 | m4, mhi = m1 | m2;      // x - nx in {0, -1}; x + nx in {nx+1, nx+2}
        l = W::sel(mlo, l + w.ld_at(o, w.x - nx, y, mlo), l);
        l = W::sel(mhi, l + w.ld_at(o, w.x + nx, y, mhi), l);
    } else {
        const auto m = m1 | m2 | m3 | m4;
        const auto src = W::seli(m1, 0, W::seli(m2, -1, W::seli(m3, nx + 1, nx + 2)));
        l = W::sel(m, l + oma * w.ld_at(o, src, y, m), l);
    }
    return l;
}

template <class W, bool PER>
SWM_HD typename W::V fold_e(const W &w, uint32_t o, int y, typename W::V l, int nx, int ny) {
    using V = typename W::V;
    const auto mx0 = w.x == (PER ? nx : 1), mx1 = w.x == (PER ? 1 : nx);
    const auto mx = mx0 | mx1;
    const auto xa = W::seli(mx0, 0, nx + 1);
    l = W::sel(mx, l + w.ld_at(o, xa, y, mx), l);
    const int ya = y == 1 ? 0 : (y == ny ? ny + 1 : -100);
    if (ya != -100) {
        V c = w.ld(o, ya);
        c = W::sel(mx, c + w.ld_at(o, xa, ya, mx), c);
        l = l + c;
    }
    return l;
}

// ------------------------------------------------------------------ loads a row ahead
//
// The marches issue the loads of their next row while the current row is computed: a load
// waited for in the row that issues it would also wait for every load issued before it.
// Each row is clamped to the rows the march reads, so every row issues the same loads.

template <class V> struct StageLoads {
    V u, v, e, H, fq, qe;      // stage input at row r; f_q and q(nx) at corner row r-1
    V lu, lv, le, mu, mv, me;  // cells of λ0 and μ_{k+1} at row r-1 (the folds add their halo)
    V xau, xav, xae, xbu, xbv, xbe;   // the first stage: cells of the extra μ's at row r-1
    V fx;                      // λFx at the row emitted, r-2 (fxg)
    V fu, fv;                  // linear dynamics: f_u at row r-1, f_v at row r-2
    V qk, qke;                 // kept q at corner row r-1, and at its column nx (periodic wrap)
    V lq, lk;                  // λ of the kept q and KEu+KEv at the row emitted, r-2
};

template <class W, bool PER, bool NEXT, int NX, int SCH = ADJ_AH>
SWM_HD StageLoads<typename W::V> stage_loads(const W &w, const StageAdjArgs<typename W::R> &A, int r, int ya, int yb) {
    using V = typename W::V;
    using R = typename W::R;
    constexpr bool LIN = (SCH & 3) == ADJ_LIN, KEEP = (SCH & ADJ_KEEP) != 0, CONT = SCH == ADJ_CONT;
    const V zero = V(R(0));
    StageLoads<V> p;
    const int rs = amin(r, yb + 1), ny = A.ny;
    if (LIN) {                 // the linear stage's transpose does not depend on the state
        p.u = p.v = p.e = p.H = p.fq = p.qe = zero;
        p.fu = w.ld_s(A.fu, amax(rs - 1, -1), A.fu_2d);
        p.fv = w.ld_s(A.fv, amax(rs - 2, -1), A.fv_2d);
    } else {
        p.u = w.ld(A.s.u, rs); p.v = w.ld(A.s.v, rs); p.e = w.ld(A.s.e, rs); p.H = w.ld(A.H, rs);
        p.fq = KEEP || CONT ? zero : w.ld_s(A.fq, rs - 1, A.fq_2d);
        p.qe = PER && !KEEP && !CONT ? w.ld_row1(A.qend, rs - 1) : zero;
        p.fu = p.fv = zero;
    }
    p.qk = p.qke = p.lq = p.lk = zero;
    if (KEEP) {
        const int yq = amin(amax(rs - 1, 0), ny), ye = amin(amax(rs - 2, ya), yb - 1);
        p.qk = w.ld(A.qk, yq);
        if (PER) p.qke = w.ld_at(A.qk, A.nx, yq, true);
        p.lq = w.ld(A.lqk, ye);
        p.lk = w.ld(A.lkk, ye);
    }
    const int yu = amin(amax(rs - 1, 1), ny), yv = amin(amax(rs - 1, 1), amax(ny - 1, 1));
    p.lu = w.ld(A.l0.u, yu); p.lv = w.ld(A.l0.v, yv); p.le = w.ld(A.l0.e, yu);
    p.mu = p.mv = p.me = zero;
    if (NEXT) { p.mu = w.ld(A.mu_in.u, yu); p.mv = w.ld(A.mu_in.v, yv); p.me = w.ld(A.mu_in.e, yu); }
    p.xau = p.xav = p.xae = p.xbu = p.xbv = p.xbe = zero;
    if (NX >= 1) { p.xau = w.ld(A.mu_x[0].u, yu); p.xav = w.ld(A.mu_x[0].v, yv); p.xae = w.ld(A.mu_x[0].e, yu); }
    if (NX >= 2) { p.xbu = w.ld(A.mu_x[1].u, yu); p.xbv = w.ld(A.mu_x[1].v, yv); p.xbe = w.ld(A.mu_x[1].e, yu); }
    const int ye = amin(amax(rs - 2, ya), yb - 1);
    p.fx = A.fxg && !CONT ? w.ld(A.lfx, ye) : zero;
    return p;
}

// ------------------------------------------------------------------ stage adjoint
//
// Forward (rhs.jl:131-324, PV_advection.jl, continuity.jl), staggered cells (x, y):
//   h = η + H;  U = u·½(h(x+1)+h);  V = v·½(h(y+1)+h);  hq = ¼(h+h(x+1)) + ¼(h(y+1)+h(x+1,y+1))
//   q = (f + v(x+1)-v - (u(y+1)-u)) / hq;  KEu = ½(u²+u(x-1)²);  KEv = ½(v²+v(y-1)²)
//   p = ½(KEu+KEv) + g·η;  qα..qδ = (1/12)·(three q's);  qhv, qhu (Arakawa-Hsu)
//   du = qhv - (p(x+1)-p) + F;  dv = -(qhu + (p(y+1)-p)) + F;  dη = -((U-U(x-1)) + (V-V(y-1))) (- γη)
// Reverse, with gu = λdu, gv = λdv, ge = λdη:
//   λp = (gu - gu(x-1)) + (gv - gv(y-1))
//   λU = (ge(x+1) - ge) - (gv(x+1)·qα + gv·qβ + gv(y-1)·qγ + gv(x+1,y-1)·qδ)
//   λV = (ge(y+1) - ge) + gu(x-1)·qα(x-1) + gu·qβ + gu(y+1)·qγ(y+1) + gu(x-1,y+1)·qδ(x-1,y+1)
//   λqα = gu·V(x+1) - gv(x+1)·U;  λqβ = gu·V - gv·U;  λqγ = gu·V(y-1) - gv(y-1)·U
//   λqδ = gu·V(x+1,y-1) - gv(x+1,y-1)·U
//   λq = (1/12)·Σ of the λqα..δ of the weights that read q;  λnum = λq/hq;  λhq = -λnum·q
//   λh = ½(λU·u + (λU·u)(x-1)) + ½(λV·v + (λV·v)(y-1)) + ¼·Σ λhq of the four corners of h
//   λu = λU·hu + u·½(λp + λp(x+1)) + (λnum - λnum(y-1))
//   λv = λV·hv + v·½(λp + λp(y+1)) + (λnum(x-1) - λnum)
//   λη = λh + g·λp (- γ·ge)
// The march reads stage-input row r and tendency-adjoint row r-1 and emits row r-2.
//
// Stage bookkeeping (classic RK, time_integration.jl:55-78): with λ0 the adjoint of u0
// after the stages and λs_{k+1} the folded μ_{k+1},
//   λd_k = a_k·λ0 + b_k·λs_{k+1},  μ_k = (∂R/∂s_k)ᵀ λd_k,
// and the adjoint of the step's input A_1 + μ_1, A_1 = λ0 + Σ_{k>1} λs_k summed as the
// recursion A_k = A_{k+1} + λs_{k+1} from A_RKo = λ0 would (((λ0 + λs_4) + λs_3) + λs_2).
// NEXT: the stage has a successor (k < RKo).  PRE: the first stage, which forms the step
// input's adjoint from λ0, its own μ_2 and NX = RKo - 2 more μ's (mu_x, oldest first).
//
// SAD: Sadourny's enstrophy-conserving PV fluxes (PV_advection.jl:31-59) in place of
// Arakawa-Hsu's, k12 = ½ (the weight of Iy!(q_u,q) and Ix!(q_v,q)):
//   qhv = Qu·Vu, Qu = ½(q + q(y-1)) at the u point, Vu = ¼(V(y-1) + V(x+1,y-1)) + ¼(V + V(x+1))
//   qhu = Qv·Uv, Qv = ½(q + q(x-1)) at the v point, Uv = ¼(U(x-1) + U) + ¼(U(x-1,y+1) + U(y+1))
// (no column wraps: q_u and q_v read no q[end,·]).  Reverse, with λqhv = gu, λqhu = -gv:
//   λQu = gu·Vu, λVu = gu·Qu, λQv = -gv·Uv, λUv = -gv·Qv, and each of them gathered back
//   onto the q, V, U it read.
//
// LIN: the linear dynamics (rhs_linear!, rhs.jl:55-86), with f_u, f_v on the u- and v-grid:
//   du = f_u·(¼(v(y-1) + v(x+1,y-1)) + ¼(v + v(x+1))) - (gη(x+1) - gη) + F
//   dv = -(f_v·(¼(u(x-1) + u) + ¼(u(x-1,y+1) + u(y+1))) + (gη(y+1) - gη)) + F
//   dη = 0 under the classic RK scheme (no mass fluxes are formed, continuity.jl:83-85),
//   plus the relaxation's -γη.  The transpose is independent of the state:
//   λu = -¼·((f_v·gv + (f_v·gv)(x+1)) + ((f_v·gv)(y-1) + (f_v·gv)(x+1,y-1)))
//   λv = ¼·((f_u·gu + (f_u·gu)(x-1)) + ((f_u·gu)(y+1) + (f_u·gu)(x-1,y+1)))
//   λη = g·λp (- γ·ge), λp as above.
//
// CONT: gu = gv = 0 (and no PV): λU = ge(x+1) - ge, λV = ge(y+1) - ge, λu = λU·hu, λv = λV·hv,
//   λη = λh (- γ·ge): the transpose of dη = -((U-U(x-1)) + (V-V(y-1))) (- γη) alone.
template <class W, bool PER, bool NEXT, int NX, bool PRE, int SCH = ADJ_AH>
SWM_HD void stage_adjoint(W &w, const StageAdjArgs<typename W::R> &A, int ya, int yb) {
    using R = typename W::R;
    using V = typename W::V;
    using M = typename W::M;
    constexpr bool SAD = (SCH & 3) == ADJ_SAD, LIN = (SCH & 3) == ADJ_LIN, KEEP = (SCH & ADJ_KEEP) != 0;
    constexpr bool CONT = SCH == ADJ_CONT;
    static_assert(!(LIN && KEEP), "the linear dynamics keep nothing");
    static_assert(!(CONT && NEXT), "the continuity form is not part of a classic RK stage");
    static_assert(NX == 0 || (NEXT && PRE), "extra μ's are summed by the first stage only");
    const int nx = A.nx, ny = A.ny;
    const auto x = w.x;
    const M ucol = PER ? M((x >= 0) & (x <= nx - 1)) : M((x >= 1) & (x <= nx - 1));
    const M vcol = (x >= 1) & (x <= nx);
    const M qcol = (x >= 0) & (x <= nx);
    const M qfix = PER && !SAD ? M(x == 0) : M(x == -1000);
    const M qdrop = PER && !SAD ? M(x == -1) : M(x == -1000);   // q(-1) is not read: qβ(0), qγ(0) read q(nx)
    const M st_u = w.own & (PER ? M((x >= -2) & (x <= nx + 1)) : ucol);
    const M st_v = w.own & (x >= -1) & (x <= nx + 2);
    const M st_e = w.own & (x >= 0) & (x <= nx + 1);
    const R half = R(0.5), quarter = R(0.25), k12 = A.k12, g = A.g;
    const V zero = V(R(0));

    // forward rows kept: inputs, thicknesses, fluxes, PV and the Arakawa-Hsu weights
    V u_1 = zero, u_2 = zero, v_1 = zero, v_2 = zero, h_1 = zero, hx_1 = zero, hu_1 = zero, hu_2 = zero, hv_2 = zero;
    V U_1 = zero, V_2 = zero, q_2 = zero, qxp_2 = zero, qxm_2 = zero, qend_2 = zero, hq_2 = zero, qa_2 = zero, qb_2 = zero;
    // adjoint rows kept
    V gu_2 = zero, gv_2 = zero, ge_2 = zero;
    V lp_2 = zero, lU_2 = zero, la_2 = zero, lb_2 = zero, lg_2 = zero, ld_2 = zero, lnum_3 = zero, lhv_3 = zero, lhq_3 = zero;
    V su_2 = zero, sv_2 = zero, se_2 = zero;           // λs_{k+1} of row r-2
    V zu_2 = zero, zv_2 = zero, ze_2 = zero;           // folded λ0 of row r-2
    V xau_2 = zero, xav_2 = zero, xae_2 = zero, xbu_2 = zero, xbv_2 = zero, xbe_2 = zero;   // folded extra μ's, row r-2
    V Qu_2 = zero, Vu_2 = zero, Uv_2 = zero;           // Sadourny: Qu, Vu of u row r-2, Uv of v row r-2
    V pu_2 = zero, pv_3 = zero;                        // linear: f_u·λdu of row r-2, f_v·λdv of row r-3

    StageLoads<V> nl = stage_loads<W, PER, NEXT, NX, SCH>(w, A, ya - 2, ya, yb);
    for (int r = ya - 2; r <= yb + 1; ++r) {
        const StageLoads<V> c = nl;
        nl = stage_loads<W, PER, NEXT, NX, SCH>(w, A, r + 1, ya, yb);
        // ---- stage input row r
        const V u_0 = c.u, v_0 = c.v, e_0 = c.e, H_0 = c.H;
        const V fq_1 = c.fq;
        const V qend_1 = KEEP ? c.qke : c.qe;
        // ---- tendency adjoint row r-1
        const int yd = r - 1;
        const bool ydu = yd >= 1 && yd <= ny, ydv = yd >= 1 && yd <= ny - 1;
        V zu = zero, zv = zero, ze = zero, su = zero, sv = zero, se = zero;
        V xau = zero, xav = zero, xae = zero, xbu = zero, xbv = zero, xbe = zero;
        if (ydu) {
            zu = fold_u<W, PER>(w, A.l0.u, yd, c.lu, A.oma, nx, ny);
            ze = fold_e<W, PER>(w, A.l0.e, yd, c.le, nx, ny);
            if (NEXT) {
                su = fold_u<W, PER>(w, A.mu_in.u, yd, c.mu, A.oma, nx, ny);
                se = fold_e<W, PER>(w, A.mu_in.e, yd, c.me, nx, ny);
            }
            if (NX >= 1) {
                xau = fold_u<W, PER>(w, A.mu_x[0].u, yd, c.xau, A.oma, nx, ny);
                xae = fold_e<W, PER>(w, A.mu_x[0].e, yd, c.xae, nx, ny);
            }
            if (NX >= 2) {
                xbu = fold_u<W, PER>(w, A.mu_x[1].u, yd, c.xbu, A.oma, nx, ny);
                xbe = fold_e<W, PER>(w, A.mu_x[1].e, yd, c.xbe, nx, ny);
            }
        }
        if (ydv) {
            zv = fold_v<W, PER>(w, A.l0.v, yd, c.lv, A.oma, nx);
            if (NEXT) sv = fold_v<W, PER>(w, A.mu_in.v, yd, c.mv, A.oma, nx);
            if (NX >= 1) xav = fold_v<W, PER>(w, A.mu_x[0].v, yd, c.xav, A.oma, nx);
            if (NX >= 2) xbv = fold_v<W, PER>(w, A.mu_x[1].v, yd, c.xbv, A.oma, nx);
        }
        const V gu_1 = CONT ? zero : W::sel(ucol & ydu, NEXT ? A.a * zu + A.b * su : A.a * zu, zero);
        const V gv_1 = CONT ? zero : W::sel(vcol & ydv, NEXT ? A.a * zv + A.b * sv : A.a * zv, zero);
        const V ge_1 = W::sel(vcol & ydu, NEXT ? A.a * ze + A.b * se : A.a * ze, zero);

        V lu_2, lv_2, le_2, lp_1;
        // ---- forward: U of row r; V, hq, q and the weights of row r-1
        const V h_0 = e_0 + H_0;
        const V hx_0 = W::xp(h_0);
        const V hu_0 = half * (hx_0 + h_0);
        const V U_0 = u_0 * hu_0;
        const V hv_1 = half * (h_0 + h_1);
        const V V_1 = v_1 * hv_1;
        const V hq_1 = quarter * (h_1 + hx_1) + quarter * (h_0 + hx_0);
        const V dvdx_1 = W::xp(v_1) - v_1;
        const V dudy_1 = u_0 - u_1;
        const bool yq1 = r - 1 >= 0 && r - 1 <= ny;
        const M qv_1 = qcol & yq1;
        const V q_1 = CONT ? zero
                    : KEEP ? W::sel(qv_1, c.qk, zero)
                           : W::sel(qv_1, (fq_1 + dvdx_1 - dudy_1) / W::sel(qv_1, hq_1, V(R(1))), zero);
        const V qxp_1 = W::xp(q_1);
        const V qxm_1 = PER && !SAD ? W::sel(qfix, qend_1, W::xm(q_1)) : W::xm(q_1);
        V qa_1 = zero, qb_1 = zero, qg_1 = zero, qd_1 = zero, la_1 = zero, lb_1 = zero, lg_1 = zero, ld_1 = zero;
        V lU_1, lq_2, lV_2, Qu_1 = zero, Vu_1 = zero, Uv_1 = zero;
        lp_1 = (gu_1 - W::xm(gu_1)) + (gv_1 - gv_2);
        if constexpr (!SAD) {
            qa_1 = k12 * (q_2 + q_1 + qxp_1);
            qb_1 = k12 * (q_2 + q_1 + qxm_1);
            qg_1 = PER ? W::sel(qfix, k12 * (q_2 + q_1 + qend_2), k12 * (q_2 + qxm_2 + q_1)) : k12 * (q_2 + qxm_2 + q_1);
            qd_1 = k12 * (q_2 + q_1 + qxp_2);

            // ---- adjoint of row r-1: Bernoulli potential, U, the Arakawa-Hsu weights
            const V gvx_1 = W::xp(gv_1), gvx_2 = W::xp(gv_2);
            lU_1 = (W::xp(ge_1) - ge_1) - (((gvx_1 * qa_1 + gv_1 * qb_1) + gv_2 * qg_1) + gvx_2 * qd_1);
            la_1 = gu_1 * W::xp(V_1) - gvx_1 * U_1;
            lb_1 = gu_1 * V_1 - gv_1 * U_1;
            lg_1 = gu_1 * V_2 - gv_2 * U_1;
            ld_1 = gu_1 * W::xp(V_2) - gvx_2 * U_1;

            // ---- adjoint of row r-2: PV, V
            const V xlb = PER ? W::sel(qdrop, zero, W::xp(lb_2)) : W::xp(lb_2);
            const V xlg = PER ? W::sel(qdrop, zero, W::xp(lg_1)) : W::xp(lg_1);
            lq_2 = k12 * ((((la_1 + la_2) + W::xm(la_2)) + ((lb_1 + lb_2) + xlb)) +
                          (((lg_1 + xlg) + lg_2) + ((ld_1 + ld_2) + W::xm(ld_1))));
            lV_2 = (ge_1 - ge_2) + (((W::xm(gu_2 * qa_2) + gu_2 * qb_2) + gu_1 * qg_1) + W::xm(gu_1 * qd_1));
        } else {
            // Sadourny: qhv of u row r-1 reads q rows r-1, r-2 and V rows r-2, r-1; qhu of v row
            // r-1 reads q row r-1 and U rows r-1, r (and v row r-2 likewise, one row down)
            Qu_1 = k12 * (q_1 + q_2);
            Vu_1 = quarter * (V_2 + W::xp(V_2)) + quarter * (V_1 + W::xp(V_1));
            Uv_1 = quarter * (W::xm(U_1) + U_1) + quarter * (W::xm(U_0) + U_0);
            const V Qv_1 = k12 * (q_1 + qxm_1), Qv_2 = k12 * (q_2 + qxm_2);
            const V lUv_1 = -(gv_1 * Qv_1), lUv_2 = -(gv_2 * Qv_2);
            lU_1 = (W::xp(ge_1) - ge_1) + (quarter * (lUv_1 + W::xp(lUv_1)) + quarter * (lUv_2 + W::xp(lUv_2)));
            const V lVu_1 = gu_1 * Qu_1, lVu_2 = gu_2 * Qu_2;
            lV_2 = (ge_1 - ge_2) + (quarter * (lVu_1 + W::xm(lVu_1)) + quarter * (lVu_2 + W::xm(lVu_2)));
            const V lQv_2 = -(gv_2 * Uv_2);
            lq_2 = k12 * ((gu_1 * Vu_1 + gu_2 * Vu_2) + (lQv_2 + W::xp(lQv_2)));
        }

        const V lq_2_keep = lq_2;
        // ---- adjoint of row r-2: thickness, inputs
        const bool yq2 = r - 2 >= 0 && r - 2 <= ny;
        const M qv_2 = qcol & yq2;
        // (kept q: λq goes to the kept field, not through PV! to the state)
        const V lnum_2 = KEEP ? zero : W::sel(qv_2, lq_2 / W::sel(qv_2, hq_2, V(R(1))), zero);
        const V lhq_2 = -(lnum_2 * q_2);
        const V lhu_2 = lU_2 * u_2;
        const V lhv_2 = lV_2 * v_2;
        const V lh_2 = half * (lhu_2 + W::xm(lhu_2)) + half * (lhv_2 + lhv_3) +
                       quarter * ((lhq_2 + W::xm(lhq_2)) + (lhq_3 + W::xm(lhq_3)));
        if constexpr (KEEP) {       // KEu+KEv kept: the Bernoulli potential reads no velocity
            lu_2 = lU_2 * hu_2;
            lv_2 = lV_2 * hv_2;
        } else {
            lu_2 = (lU_2 * hu_2 + u_2 * (half * (lp_2 + W::xp(lp_2)))) + (lnum_2 - lnum_3);
            lv_2 = (lV_2 * hv_2 + v_2 * (half * (lp_2 + lp_1))) + (W::xm(lnum_2) - lnum_2);
        }
        le_2 = lh_2 + g * lp_2;
        if (A.cont == 1) le_2 = le_2 - A.gamma * ge_2;          // surface relaxation: dη += γ(η_ref - η)
        if constexpr (LIN) {
            // the linear dynamics' transpose (see above): f_u of rows r-1 and r-2, f_v of rows
            // r-2 and r-3, the tendency adjoint of rows r-1 .. r-3
            // (each u / v point's own f times its cotangent, then the four points summed)
            const V pu_1 = c.fu * gu_1, pv_2 = c.fv * gv_2;
            lv_2 = quarter * (pu_2 + W::xm(pu_2)) + quarter * (pu_1 + W::xm(pu_1));
            lu_2 = -(quarter * (pv_2 + W::xp(pv_2)) + quarter * (pv_3 + W::xp(pv_3)));
            le_2 = g * lp_2;
            if (A.cont == 1) le_2 = le_2 - A.gamma * ge_2;      // surface relaxation: dη += γ(η_ref - η)
            pu_2 = pu_1;
            pv_3 = pv_2;
        }

        // ---- emit row y = r-2
        const int y = r - 2;
        if (y >= ya && y < yb) {
            const bool yu_in = y >= 1 && y <= ny, yv_in = y >= 1 && y <= ny - 1;
            // every cell a later fold reads is written (u rows -1, ny+2 are not read by rhs!: μ = 0)
            const bool yu_st = PER ? yu_in : (y >= -1 && y <= ny + 2);
            const bool yv_st = yv_in, ye_st = y >= 0 && y <= ny + 1;
            const M iu = ucol & yu_in, iv = vcol & yv_in, ie = vcol & yu_in;
            if (NEXT && PRE) {
                V au = zu_2, av = zv_2, ae = ze_2;              // A_RKo = λ0, then the recursion
                if (NX >= 1) { au = au + xau_2; av = av + xav_2; ae = ae + xae_2; }
                if (NX >= 2) { au = au + xbu_2; av = av + xbv_2; ae = ae + xbe_2; }
                const V Au = au + su_2, Av = av + sv_2, Ae = ae + se_2;     // A_1 on interior cells
                w.st(A.a_out.u, y, W::sel(iu, Au + lu_2, lu_2), st_u & yu_st);
                w.st(A.a_out.v, y, W::sel(iv, Av + lv_2, lv_2), st_v & yv_st);
                w.st(A.a_out.e, y, W::sel(ie, Ae + le_2, le_2), st_e & ye_st);
            } else {
                w.st(A.mu_out.u, y, lu_2, st_u & yu_st);
                w.st(A.mu_out.v, y, lv_2, st_v & yv_st);
                w.st(A.mu_out.e, y, le_2, st_e & ye_st);
            }
            // periodic: λ of q(nx, y) from qβ(0, y) and qγ(0, y+1)
            if (PER && !SAD && !LIN && !CONT && y >= 0 && y <= ny) w.st_row1(A.lqend, y, k12 * (lb_2 + lg_1), w.own & (x == 0));
            // the wind stress enters du as Tprog(Fxt·Fx) (rhs.jl:242): λFx += Fxt·λdu, u cells
            if (!CONT && A.fxg) w.st(A.lfx, y, c.fx + A.fxt * gu_2, w.own & iu);
            if constexpr (KEEP) {   // λ of the kept q (corner row y) and KEu+KEv (p = ½(KEu+KEv) + gη)
                if (y >= 0 && y <= ny) w.st(A.lqk, y, c.lq + W::sel(qcol, lq_2_keep, zero), w.own & qcol);
                if (y >= 0 && y <= ny + 1) w.st(A.lkk, y, c.lk + half * lp_2, w.own & (x >= 0) & (x <= nx + 1));
            }
        }

        // ---- advance the rows
        u_2 = u_1; u_1 = u_0; v_2 = v_1; v_1 = v_0; h_1 = h_0; hx_1 = hx_0; hu_2 = hu_1; hu_1 = hu_0; hv_2 = hv_1;
        U_1 = U_0; V_2 = V_1; q_2 = q_1; qxp_2 = qxp_1; qxm_2 = qxm_1; qend_2 = qend_1; hq_2 = hq_1; qa_2 = qa_1; qb_2 = qb_1;
        gu_2 = gu_1; gv_2 = gv_1; ge_2 = ge_1;
        lp_2 = lp_1; lU_2 = lU_1; la_2 = la_1; lb_2 = lb_1; lg_2 = lg_1; ld_2 = ld_1;
        lnum_3 = lnum_2; lhv_3 = lhv_2; lhq_3 = lhq_2;
        su_2 = su; sv_2 = sv; se_2 = se;
        zu_2 = zu; zv_2 = zv; ze_2 = ze;
        if constexpr (NX >= 1) { xau_2 = xau; xav_2 = xav; xae_2 = xae; }
        if constexpr (NX >= 2) { xbu_2 = xbu; xbv_2 = xbv; xbe_2 = xbe; }
        if constexpr (SAD) { Qu_2 = Qu_1; Vu_2 = Vu_1; Uv_2 = Uv_1; }
    }
}

// Periodic only: the contribution of λq(nx, y) (qend) to μ at the cells q(nx, y) reads:
// u(nx, y..y+1), v(nx..nx+1, y), η(nx..nx+1, y..y+1) (PV!, PV_advection.jl:2-18).  One
// thread per row yy = 0 .. ny+1 gathers corners yy-1 and yy.  C is the cell accessor.
template <class R, class C>
SWM_HD void qend_adjoint_row(const C &c, const StageAdjArgs<R> &A, const UVE &mu, int yy) {
    const int nx = A.nx, ny = A.ny;
    const R quarter = R(0.25);
    R ln[2] = {R(0), R(0)}, lh[2] = {R(0), R(0)};     // corners yy-1, yy
    for (int k = 0; k < 2; ++k) {
        const int y = yy - 1 + k;
        if (y < 0 || y > ny) continue;
        const R h00 = c.at(A.s.e, nx, y) + c.at(A.H, nx, y), h10 = c.at(A.s.e, nx + 1, y) + c.at(A.H, nx + 1, y);
        const R h01 = c.at(A.s.e, nx, y + 1) + c.at(A.H, nx, y + 1);
        const R h11 = c.at(A.s.e, nx + 1, y + 1) + c.at(A.H, nx + 1, y + 1);
        const R hq = quarter * (h00 + h10) + quarter * (h01 + h11);
        const R q = c.row1(A.qend, y);
        ln[k] = c.row1(A.lqend, y) / hq;
        lh[k] = -(ln[k] * q);
    }
    // dudy(nx, y) = u(nx, y+1) - u(nx, y): λu(nx, yy) += λnum(yy) - λnum(yy-1)
    if (yy >= 1 && yy <= ny) c.at(mu.u, nx, yy) = c.at(mu.u, nx, yy) + (ln[1] - ln[0]);
    // dvdx(nx, y) = v(nx+1, y) - v(nx, y)
    if (yy <= ny - 1 && yy >= 1) {
        c.at(mu.v, nx + 1, yy) = c.at(mu.v, nx + 1, yy) + ln[1];
        c.at(mu.v, nx, yy) = c.at(mu.v, nx, yy) - ln[1];
    }
    // hq: ¼ to each of the four cells of the corner
    const R le = quarter * (lh[0] + lh[1]);
    c.at(mu.e, nx, yy) = c.at(mu.e, nx, yy) + le;
    c.at(mu.e, nx + 1, yy) = c.at(mu.e, nx + 1, yy) + le;
}

// Periodic with kept q: the λ of q(nx, y) the wrap gathered (lqend) belongs to the kept
// field's corner (nx, y).  One thread per corner row.
template <class R, class C> SWM_HD void lqend_keep_row(const C &c, const StageAdjArgs<R> &A, int y) {
    c.at(A.lqk, A.nx, y) = c.at(A.lqk, A.nx, y) + c.row1(A.lqend, y);
}

// ------------------------------------------------------------------ kept q and KEu+KEv
//
// nstep_advcor > 0 (time_integration.jl:20-31, 208-211): q and KEu+KEv are formed from a
// state (u, v, η) -- the initial state, the state after the stages of every nstep_advcor-th
// step -- and read by the stages of the following steps; the quadratic drag forms KEu+KEv
// again from the state after the stages of its steps (bottom_drag.jl:25-38).  Forward, at
// the state (T points (x, y), corners (x, y)):
//   KE = ½(u² + u(x-1)²) + ½(v² + v(y-1)²)
//   q = (f + (v(x+1) - v) - (u(y+1) - u)) / hq,  hq = ¼(h + h(x+1)) + ¼(h(y+1) + h(x+1,y+1))
// Transpose, with λk, λq the accumulated cotangents of the kept fields (zero outside their
// cells), λn = λq/hq, λhq = -λn·q:
//   λu += u·(λk + λk(x+1)) + (λn - λn(y-1));  λv += v·(λk + λk(y+1)) + (λn(x-1) - λn)
//   λη += ¼·(λhq + λhq(x-1) + λhq(y-1) + λhq(x-1,y-1))
// added to the (halo unfolded) adjoint of the state.  wq: q was formed too (else KE alone).
struct AdvcorAdjArgs {
    UVE s;                 // the state the kept fields were formed from
    uint32_t H, fq;
    int fq_2d;
    uint32_t lqk, lkk;     // λ of the kept q and KEu+KEv
    UVE out;               // the adjoint of the state, added to
    int wq, nx, ny;
};

template <class R, class C> SWM_HD void advcor_adjoint_cell(const C &c, const AdvcorAdjArgs &A, int x, int y) {
    const int nx = A.nx, ny = A.ny;
    const R quarter = R(0.25), zero = R(0);
    auto lk = [&](int xx, int yy) { return (xx >= 0 && xx <= nx + 1 && yy >= 0 && yy <= ny + 1) ? c.at(A.lkk, xx, yy) : zero; };
    auto cq = [&](int xx, int yy) { return xx >= 0 && xx <= nx && yy >= 0 && yy <= ny; };
    auto hq = [&](int xx, int yy) {
        const R h00 = c.at(A.s.e, xx, yy) + c.at(A.H, xx, yy), h10 = c.at(A.s.e, xx + 1, yy) + c.at(A.H, xx + 1, yy);
        const R h01 = c.at(A.s.e, xx, yy + 1) + c.at(A.H, xx, yy + 1);
        const R h11 = c.at(A.s.e, xx + 1, yy + 1) + c.at(A.H, xx + 1, yy + 1);
        return quarter * (h10 + h00) + quarter * (h11 + h01);
    };
    auto ln = [&](int xx, int yy) { return cq(xx, yy) ? c.at(A.lqk, xx, yy) / hq(xx, yy) : zero; };
    auto lh = [&](int xx, int yy) {
        if (!cq(xx, yy)) return zero;
        const R d = hq(xx, yy);
        const R f = A.fq_2d ? c.at(A.fq, xx, yy) : c.row1(A.fq, yy);
        const R q = (f + (c.at(A.s.v, xx + 1, yy) - c.at(A.s.v, xx, yy)) - (c.at(A.s.u, xx, yy + 1) - c.at(A.s.u, xx, yy))) / d;
        return -((c.at(A.lqk, xx, yy) / d) * q);
    };
    R du = c.at(A.s.u, x, y) * (lk(x, y) + lk(x + 1, y));
    R dv = c.at(A.s.v, x, y) * (lk(x, y) + lk(x, y + 1));
    R de = zero;
    if (A.wq) {
        du = du + (ln(x, y) - ln(x, y - 1));
        dv = dv + (ln(x - 1, y) - ln(x, y));
        de = quarter * ((lh(x, y) + lh(x - 1, y)) + (lh(x, y - 1) + lh(x - 1, y - 1)));
    }
    c.at(A.out.u, x, y) = c.at(A.out.u, x, y) + du;
    c.at(A.out.v, x, y) = c.at(A.out.v, x, y) + dv;
    c.at(A.out.e, x, y) = c.at(A.out.e, x, y) + de;
}

// q at column nx of every corner row y (PV!, the value the periodic qβ/qγ wrap reads)
template <class R, class C> SWM_HD void qend_row(const C &c, const StageAdjArgs<R> &A, int y) {
    const int x = A.nx;
    const R quarter = R(0.25);
    const R h00 = c.at(A.s.e, x, y) + c.at(A.H, x, y), h10 = c.at(A.s.e, x + 1, y) + c.at(A.H, x + 1, y);
    const R h01 = c.at(A.s.e, x, y + 1) + c.at(A.H, x, y + 1), h11 = c.at(A.s.e, x + 1, y + 1) + c.at(A.H, x + 1, y + 1);
    const R hq = quarter * (h00 + h10) + quarter * (h01 + h11);
    const R dvdx = c.at(A.s.v, x + 1, y) - c.at(A.s.v, x, y);
    const R dudy = c.at(A.s.u, x, y + 1) - c.at(A.s.u, x, y);
    const R f = A.fq_2d ? c.at(A.fq, x, y) : c.row1(A.fq, y);
    c.row1(A.qend, y) = (f + dvdx - dudy) / hq;
}

// ------------------------------------------------------------------ dissipation adjoint
//
// Forward (per the dissipation kernel), with interior increments u' = u + Δt_d·(Bu + LLu1 +
// LLu2), v' likewise:
//   Lu = ∇²u, Lv = ∇²v;  S11 = ν·(Lu - Lu(x-1)), S22 = ν·(Lv - Lv(y-1)) (T-points),
//   S12 = ν_q·(Lu(y+1) - Lu), S21 = ν_q·(Lv(x+1) - Lv) (corners);
//   LLu1 = S11(x+1) - S11, LLu2 = S12 - S12(y-1), LLv1 = S21 - S21(x-1), LLv2 = S22(y+1) - S22
//   Smagorinsky: ν = cS·√(DS + DT), DT = ((u-u(x-1)) + (v-v(y-1)))², DS = ¼Σ DSq of the four
//   corners, DSq = ((u(y+1)-u) + (v(x+1)-v))², ν_q = ¼Σ ν of the four cells
//   quadratic drag: Bu = cD·½(sk(x+1)+sk)·u / ½(h(x+1)+h), sk = √(½(u²+u(x-1)²) + ½(v²+v(y-1)²))
// The march reads state row r and the adjoint after the dissipation at row r-1, and emits
// row r-3.
template <class W, bool PER, int DRAG, bool SMAG>
SWM_HD void diss_adjoint(W &w, const DissAdjArgs<typename W::R> &A, int ya, int yb) {
    using R = typename W::R;
    using V = typename W::V;
    using M = typename W::M;
    const int nx = A.nx, ny = A.ny;
    const auto x = w.x;
    const M ucol = PER ? M((x >= 0) & (x <= nx - 1)) : M((x >= 1) & (x <= nx - 1));
    const M vcol = (x >= 1) & (x <= nx);
    const M st_u = w.own & (PER ? M((x >= -2) & (x <= nx + 1)) : ucol);
    const M st_v = w.own & (x >= -1) & (x <= nx + 2);
    const M st_e = w.own & (x >= 0) & (x <= nx + 1);
    const R half = R(0.5), quarter = R(0.25), two = R(2), cD = A.cD, cS = A.cS, dtd = A.dtd;
    const V zero = V(R(0));

    V u_1 = zero, u_2 = zero, u_3 = zero, uxm_1 = zero, v_1 = zero, v_2 = zero, v_3 = zero, h_1 = zero, h_2 = zero;
    V sk_1 = zero, sk_2 = zero, sk_3 = zero;
    V Lu_1 = zero, Lv_1 = zero, Lvxp_2 = zero;
    V DT_1 = zero, DSq_2 = zero, DSqxm_2 = zero, nu_2 = zero, nuxp_2 = zero;
    V s0_1 = zero, s0_2 = zero, s1_2 = zero, s1_3 = zero, snu_2 = zero;
    V gu_2 = zero, gv_2 = zero, lin_u_2 = zero, lin_u_3 = zero, lin_v_2 = zero, lin_v_3 = zero;
    V ldLudx_2 = zero, ldLudy_3 = zero, ldLvdy_2 = zero;
    V lnuS_2 = zero, lnuq_3 = zero, lz_3 = zero, ls0_3 = zero, ls1_4 = zero;
    V lLu_3 = zero, lLu_4 = zero, lLv_3 = zero, lLv_4 = zero;
    V lKE_3 = zero, lskv_3 = zero, lhu_3 = zero, lhv_3 = zero, lhv_4 = zero, bdu_3 = zero, bdv_3 = zero;

    // loads a row ahead (rows clamped to the rows the march reads): the state at row r, the
    // cells of λ after the dissipation at row r-1 and η's at the row emitted, r-3
    struct Loads {
        V u, v, e, H, lu, lv, le;
    };
    auto loads = [&](int r) {
        Loads p;
        const int rs = amin(r, yb + 2);
        p.u = w.ld(A.s.u, rs); p.v = w.ld(A.s.v, rs);
        p.e = p.H = zero;
        if (DRAG == 0) { p.e = w.ld(A.s.e, rs); p.H = w.ld(A.H, rs); }
        p.lu = w.ld(A.lin.u, amin(amax(rs - 1, 1), ny));
        p.lv = w.ld(A.lin.v, amin(amax(rs - 1, 1), amax(ny - 1, 1)));
        p.le = w.ld(A.lin.e, amin(amax(rs - 3, ya), yb - 1));
        return p;
    };
    Loads nl = loads(ya - 3);
    for (int r = ya - 3; r <= yb + 2; ++r) {
        const Loads c = nl;
        nl = loads(r + 1);
        // ---- state row r
        const V u_0 = c.u, v_0 = c.v;
        const V e_0 = c.e, H_0 = c.H;
        // ---- adjoint after the dissipation, row r-1 (interior increments only)
        const int yi = r - 1;
        const bool yiu = yi >= 1 && yi <= ny, yiv = yi >= 1 && yi <= ny - 1;
        const V lin_u_1 = yiu ? W::sel(ucol, fold_u<W, PER>(w, A.lin.u, yi, c.lu, A.oma, nx, ny), zero) : zero;
        const V lin_v_1 = yiv ? W::sel(vcol, fold_v<W, PER>(w, A.lin.v, yi, c.lv, A.oma, nx), zero) : zero;
        const V gu_1 = dtd * lin_u_1, gv_1 = dtd * lin_v_1;

        // ---- forward (as the dissipation kernel): row r quantities, Laplacians of row r-1
        const V uxm_0 = W::xm(u_0);
        V h_0 = zero, sk_0 = zero;
        if (DRAG == 0) {
            h_0 = e_0 + H_0;
            const V KEu_0 = half * (u_0 * u_0 + uxm_0 * uxm_0);
            const V KEv_0 = half * (v_0 * v_0 + v_1 * v_1);
            sk_0 = W::vsqrt(KEu_0 + KEv_0);
        }
        const V Lu_1n = ((uxm_1 - u_1) + (W::xp(u_1) - u_1)) + ((u_2 - u_1) + (u_0 - u_1));
        const V Lv_1n = ((W::xm(v_1) - v_1) + (W::xp(v_1) - v_1)) + ((v_2 - v_1) + (v_0 - v_1));
        V nu_1 = zero, nuq_2 = zero, DT_0 = zero, s0_0 = zero, DSq_1 = zero, DSqxm_1 = zero, s1_1 = zero, snu_1 = zero,
          nuxp_1 = zero;
        if (SMAG) {
            s0_0 = (u_0 - uxm_0) + (v_0 - v_1);
            DT_0 = s0_0 * s0_0;
            s1_1 = (u_0 - u_1) + (W::xp(v_1) - v_1);
            DSq_1 = s1_1 * s1_1;
            DSqxm_1 = W::xm(DSq_1);
            const V DS_1 = quarter * (DSqxm_2 + DSq_2) + quarter * (DSqxm_1 + DSq_1);
            snu_1 = W::vsqrt(DS_1 + DT_1);
            nu_1 = cS * snu_1;
            nuxp_1 = W::xp(nu_1);
            nuq_2 = quarter * (nu_2 + nuxp_2) + quarter * (nu_1 + nuxp_1);
        } else {
            nu_1 = V(A.nuB);
            nuq_2 = V(A.nuB);
        }
        const V dLudx_1 = Lu_1n - W::xm(Lu_1n);             // T, row r-1
        const V dLvdy_1 = Lv_1n - Lv_1;                     // T, row r-1 (Lv_1 holds row r-2 here)
        const V dLudy_2 = Lu_1n - Lu_1;                     // corner, row r-2
        const V Lvxp_1 = W::xp(Lv_1n);
        const V dLvdx_2 = Lvxp_2 - Lv_1;                    // corner, row r-2 (Lvxp_2 = Lv(x+1, r-2))

        // ---- adjoint of the viscous tensor: S11, S22 at row r-1, S12, S21 at corner row r-2
        const V lS11_1 = W::xm(gu_1) - gu_1;
        const V lS22_1 = gv_2 - gv_1;
        const V lS12_2 = gu_2 - gu_1;
        const V lS21_2 = gv_2 - W::xp(gv_2);
        const V ldLudx_1 = lS11_1 * nu_1;
        const V ldLvdy_1 = lS22_1 * nu_1;
        const V ldLudy_2 = lS12_2 * nuq_2;
        const V ldLvdx_2 = lS21_2 * nuq_2;
        // λLu (u-points) and λLv (v-points) of row r-2
        const V lLu_2 = (ldLudx_2 - W::xp(ldLudx_2)) + (ldLudy_3 - ldLudy_2);
        const V lLv_2 = (W::xm(ldLvdx_2) - ldLvdx_2) + (ldLvdy_2 - ldLvdy_1);

        // Smagorinsky coefficient: λν of row r-2, then λ(DS+DT), λDSq of corner row r-3
        V ls0_2 = zero, ls1_3 = zero, lz_2 = zero, lnuq_2 = zero, lnuS_1 = zero;
        if (SMAG) {
            lnuq_2 = lS12_2 * dLudy_2 + lS21_2 * dLvdx_2;
            lnuS_1 = lS11_1 * dLudx_1 + lS22_1 * dLvdy_1;
            const V lnu_2 = lnuS_2 + quarter * ((lnuq_2 + W::xm(lnuq_2)) + (lnuq_3 + W::xm(lnuq_3)));
            lz_2 = W::sel(snu_2 > zero, (lnu_2 * cS) * half / W::sel(snu_2 > zero, snu_2, V(R(1))), zero);
            ls0_2 = two * s0_2 * lz_2;
            const V lDSq_3 = quarter * ((lz_3 + W::xp(lz_3)) + (lz_2 + W::xp(lz_2)));
            ls1_3 = two * s1_3 * lDSq_3;
        }

        // quadratic drag of row r-2
        V lKE_2 = zero, lskv_2 = zero, lhu_2 = zero, lhv_2 = zero, bdu_2 = zero, bdv_2 = zero;
        if (DRAG == 0) {
            const V sku_2 = half * (W::xp(sk_2) + sk_2), hu_2 = half * (W::xp(h_2) + h_2);
            const V skv_2 = half * (sk_1 + sk_2), hv_2 = half * (h_1 + h_2);
            // the derivatives of cD·sk·u/h share 1/h: one division per point (the products
            // round differently from the forward quotient; the linearisation is the same)
            const V ihu = V(R(1)) / W::sel(hu_2 != zero, hu_2, V(R(1))), ihv = V(R(1)) / W::sel(hv_2 != zero, hv_2, V(R(1)));
            const V Bu_2 = cD * sku_2 * u_2 * ihu;
            const V Bv_2 = cD * skv_2 * v_2 * ihv;
            bdu_2 = gu_2 * ((cD * sku_2) * ihu);
            bdv_2 = gv_2 * ((cD * skv_2) * ihv);
            const V lsku_2 = gu_2 * (cD * u_2 * ihu);
            lskv_2 = gv_2 * (cD * v_2 * ihv);
            lhu_2 = -(gu_2 * Bu_2 * ihu);
            lhv_2 = -(gv_2 * Bv_2 * ihv);
            const V lsk_2 = half * (lsku_2 + W::xm(lsku_2)) + half * (lskv_2 + lskv_3);
            lKE_2 = W::sel(sk_2 > zero, lsk_2 * half / W::sel(sk_2 > zero, sk_2, V(R(1))), zero);
        } else if (DRAG == 1) {
            bdu_2 = A.rD * gu_2;
            bdv_2 = A.rD * gv_2;
        }

        // ---- emit row y = r-3
        const int y = r - 3;
        if (y >= ya && y < yb) {
            const V lapu = ((W::xm(lLu_3) - lLu_3) + (W::xp(lLu_3) - lLu_3)) + ((lLu_4 - lLu_3) + (lLu_2 - lLu_3));
            const V lapv = ((W::xm(lLv_3) - lLv_3) + (W::xp(lLv_3) - lLv_3)) + ((lLv_4 - lLv_3) + (lLv_2 - lLv_3));
            V lu = lin_u_3 + lapu, lv = lin_v_3 + lapv;
            if (SMAG) {
                lu = lu + ((ls0_3 - W::xp(ls0_3)) + (ls1_4 - ls1_3));
                lv = lv + ((ls0_3 - ls0_2) + (W::xm(ls1_3) - ls1_3));
            }
            V le = c.le;
            if (DRAG == 0) {
                lu = lu + (bdu_3 + u_3 * (lKE_3 + W::xp(lKE_3)));
                lv = lv + (bdv_3 + v_3 * (lKE_3 + lKE_2));
                le = le + (half * (lhu_3 + W::xm(lhu_3)) + half * (lhv_3 + lhv_4));
            } else if (DRAG == 1) {
                lu = lu + bdu_3;
                lv = lv + bdv_3;
            }
            const bool yu_st = PER ? (y >= 1 && y <= ny) : (y >= -1 && y <= ny + 2);
            w.st(A.out.u, y, lu, st_u & yu_st);
            w.st(A.out.v, y, lv, st_v & (y >= 1 && y <= ny - 1));
            w.st(A.out.e, y, le, st_e & (y >= 0 && y <= ny + 1));
        }

        // ---- advance
        u_3 = u_2; u_2 = u_1; u_1 = u_0; uxm_1 = uxm_0; v_3 = v_2; v_2 = v_1; v_1 = v_0; h_2 = h_1; h_1 = h_0;
        sk_3 = sk_2; sk_2 = sk_1; sk_1 = sk_0;
        Lu_1 = Lu_1n; Lv_1 = Lv_1n; Lvxp_2 = Lvxp_1;
        DT_1 = DT_0; DSq_2 = DSq_1; DSqxm_2 = DSqxm_1; nu_2 = nu_1; nuxp_2 = nuxp_1;
        s0_2 = s0_1; s0_1 = s0_0; s1_3 = s1_2; s1_2 = s1_1; snu_2 = snu_1;
        gu_2 = gu_1; gv_2 = gv_1;
        lin_u_3 = lin_u_2; lin_u_2 = lin_u_1; lin_v_3 = lin_v_2; lin_v_2 = lin_v_1;
        ldLudx_2 = ldLudx_1; ldLudy_3 = ldLudy_2; ldLvdy_2 = ldLvdy_1;
        lnuS_2 = lnuS_1; lnuq_3 = lnuq_2; lz_3 = lz_2; ls0_3 = ls0_2; ls1_4 = ls1_3;
        lLu_4 = lLu_3; lLu_3 = lLu_2; lLv_4 = lLv_3; lLv_3 = lLv_2;
        lKE_3 = lKE_2; lskv_3 = lskv_2; lhu_3 = lhu_2; lhv_4 = lhv_3; lhv_3 = lhv_2; bdu_3 = bdu_2; bdv_3 = bdv_2;
    }
}


// ------------------------------------------------------------------ per-cell kernels
//
// A cell accessor C gives R &at(origin, x, y), int &ati(origin, x, y) (an int32 stored in a
// field's element slot) and R &row1(origin, y) (a 1-D field with one value per row).

// ---- tracer (tracer_advection.jl:1-44), adjoint
//
// Forward at T-cell (x, y): u_T = ½(u + u(x-1)), v_T = ½(v + v(y-1)); departure
// (xd0, yd0) = -(½Δt_adv)·(u_T, v_T) in cells; (ui, vi) = bilinear of um_T, vm_T there;
// (xd, yd) = -Δt_adv·(ui, vi); sst' = bilinear of sst at (xd, yd) (+ consumption); the
// interpolation indices (floor, clip, wrap) are those of interp_uv! / adv_sst!.
// Adjoint, first pass (one thread per source cell): derivatives of the three bilinear
// interpolations with respect to their coordinates, λu_T, λv_T, and for each
// interpolation a record (λ of its value, x0, y0, first corner) from which the second pass
// gathers the adjoint of the interpolated field without atomics.
template <class R> struct TracerAdjArgs {
    uint32_t u, v;                  // u0, v0 after the dissipation
    uint32_t um, vm;                // mid-point velocities the step read
    uint32_t sst;                   // tracer before the step
    uint32_t lsst;                  // λ of the tracer after the step (halo unfolded)
    uint32_t lsst_out;              // λ of the tracer before the step (halo unfolded)
    uint32_t lu, lv, lum, lvm;      // accumulated λ of u0, v0, um, vm (all cells)
    uint32_t rs[4], ru[4], rv[4];   // records of the sst, um_T, vm_T interpolations: λ, x0, y0, corner
    uint32_t luT, lvT;              // λ of u_T, v_T
    uint32_t gu, gv;                // the um_T / vm_T records gathered at each corner (second pass)
    R hdu, hdv, du, dv, jSST;
    int consume, nx, ny, P, ep;
};

SWM_HD int pack_corner(int cx, int cy) { return (cx + 32) | ((cy + 32) << 16); }
SWM_HD int corner_x(int p) { return (p & 0xFFFF) - 32; }
SWM_HD int corner_y(int p) { return (p >> 16) - 32; }

SWM_HD double floor_r(double x) { return __builtin_floor(x); }
SWM_HD float floor_r(float x) { return __builtin_floorf(x); }

template <class R> struct Bilin {
    int cx, cy;        // cell of the first corner
    R x0, y0;
    bool clipx, clipy;
};

// index arithmetic of interp_uv! (tracer_advection.jl:130-173) and adv_sst! (:179-206):
// reference index k = floor(xx) + i + h1 of an array of mn points, clipped or wrapped
template <class R, bool PER>
SWM_HD Bilin<R> locate(R xx, R yy, int x, int y, int mx, int my, int xh1, int xh2, int yh1, int cx_of_k, int cy_of_l) {
    Bilin<R> b;
    R fx = floor_r(xx), fy = floor_r(yy);
    if (!(fx == fx) || fx > R(1e6) || fx < R(-1e6)) fx = R(0);     // NaN / runaway: the forward has failed
    if (!(fy == fy) || fy > R(1e6) || fy < R(-1e6)) fy = R(0);
    const long xi = (long)fx, yi = (long)fy;
    long k = xi + x + xh1, l = yi + y + yh1;
    b.x0 = xx - (R)xi;
    b.y0 = yy - (R)yi;
    b.clipx = b.clipy = false;
    if (PER) {
        if (k < 1 + xh1) k = k + mx - xh1 - xh2;
        else if (k > mx - 1) k = k - mx + xh1 + xh2;
        // a departure point farther than one period (the forward reports it as an error)
        // must still address the array
        if (k < 1 + xh1 || k > mx - 1) k = k < 1 + xh1 ? 1 + xh1 : mx - 1;
    } else {
        if (k < 1 + xh1) { k = 1 + xh1; b.x0 = R(0); b.clipx = true; }
        else if (k > mx - 1) { k = mx - 1; b.x0 = R(1); b.clipx = true; }
    }
    if (l < 1 + yh1) { l = 1 + yh1; b.y0 = R(0); b.clipy = true; }
    else if (l > my - 1) { l = my - 1; b.y0 = R(1); b.clipy = true; }
    b.cx = (int)k + cx_of_k;
    b.cy = (int)l + cy_of_l;
    return b;
}


// returns the largest distance (in cells, x and y packed as dx | dy << 16) between the
// source cell and the corners its records name
template <class R, bool PER, class C>
SWM_HD int tracer_adj_source(const C &c, const TracerAdjArgs<R> &A, int x, int y) {
    const int nx = A.nx, ny = A.ny, ep = A.ep;
    const R half = R(0.5), one = R(1), zero = R(0);
    // λ of the new tracer with its ghost copies folded (ghost_points_sst!, halo x only)
    R ls = c.at(A.lsst, x, y);
    if (x == (PER ? nx : 1)) ls = ls + c.at(A.lsst, 0, y);
    if (x == (PER ? 1 : nx)) ls = ls + c.at(A.lsst, nx + 1, y);
    if (A.consume) ls = ls + A.jSST * (-ls);            // s' = s + j·(smin - s)
    // forward departure chain
    const R uT = half * (c.at(A.u, x, y) + c.at(A.u, x - 1, y));
    const R vT = half * (c.at(A.v, x, y) + c.at(A.v, x, y - 1));
    const R xd0 = -A.hdu * uT, yd0 = -A.hdv * vT;
    // u-grid: mx = nx+2+ep, my = ny+4, halo (1+ep, 1) x, 2 y; corner cell (k-1-ep, l-2)
    const Bilin<R> bu = locate<R, PER>(xd0, yd0, x, y, nx + 2 + ep, ny + 4, 1 + ep, 1, 2, -1 - ep, -2);
    // v-grid: mx = nx+4, my = ny+2, halo 2 x, 1 y; corner cell (k-2, l-1)
    const Bilin<R> bv = locate<R, PER>(xd0, yd0, x, y, nx + 4, ny + 2, 2, 2, 1, -2, -1);
    auto umT = [&](int cx, int cy) { return half * (c.at(A.um, cx, cy) + c.at(A.um, cx - 1, cy)); };
    auto vmT = [&](int cx, int cy) { return half * (c.at(A.vm, cx, cy) + c.at(A.vm, cx, cy - 1)); };
    const R u00 = umT(bu.cx, bu.cy), u10 = umT(bu.cx + 1, bu.cy), u01 = umT(bu.cx, bu.cy + 1), u11 = umT(bu.cx + 1, bu.cy + 1);
    const R v00 = vmT(bv.cx, bv.cy), v10 = vmT(bv.cx + 1, bv.cy), v01 = vmT(bv.cx, bv.cy + 1), v11 = vmT(bv.cx + 1, bv.cy + 1);
    const R ui = u00 * (one - bu.x0) * (one - bu.y0) + u10 * bu.x0 * (one - bu.y0) + u01 * (one - bu.x0) * bu.y0 +
                 u11 * bu.x0 * bu.y0;
    const R vi = v00 * (one - bv.x0) * (one - bv.y0) + v10 * bv.x0 * (one - bv.y0) + v01 * (one - bv.x0) * bv.y0 +
                 v11 * bv.x0 * bv.y0;
    const R xd = -A.du * ui, yd = -A.dv * vi;
    // sst: m = nx+2, n = ny, halo (1, 0); corner cell (k-1, l)
    const Bilin<R> bs = locate<R, PER>(xd, yd, x, y, nx + 2, ny, 1, 1, 0, -1, 0);
    const R s00 = c.at(A.sst, bs.cx, bs.cy), s10 = c.at(A.sst, bs.cx + 1, bs.cy), s01 = c.at(A.sst, bs.cx, bs.cy + 1),
            s11 = c.at(A.sst, bs.cx + 1, bs.cy + 1);
    // reverse
    const R lxd = bs.clipx ? zero : ls * ((s10 - s00) * (one - bs.y0) + (s11 - s01) * bs.y0);
    const R lyd = bs.clipy ? zero : ls * ((s01 - s00) * (one - bs.x0) + (s11 - s10) * bs.x0);
    const R lui = -A.du * lxd, lvi = -A.dv * lyd;
    R lxd0 = zero, lyd0 = zero;
    if (!bu.clipx) lxd0 = lxd0 + lui * ((u10 - u00) * (one - bu.y0) + (u11 - u01) * bu.y0);
    if (!bu.clipy) lyd0 = lyd0 + lui * ((u01 - u00) * (one - bu.x0) + (u11 - u10) * bu.x0);
    if (!bv.clipx) lxd0 = lxd0 + lvi * ((v10 - v00) * (one - bv.y0) + (v11 - v01) * bv.y0);
    if (!bv.clipy) lyd0 = lyd0 + lvi * ((v01 - v00) * (one - bv.x0) + (v11 - v10) * bv.x0);
    c.at(A.luT, x, y) = -A.hdu * lxd0;
    c.at(A.lvT, x, y) = -A.hdv * lyd0;
    c.at(A.rs[0], x, y) = ls;  c.at(A.rs[1], x, y) = bs.x0; c.at(A.rs[2], x, y) = bs.y0; c.ati(A.rs[3], x, y) = pack_corner(bs.cx, bs.cy);
    c.at(A.ru[0], x, y) = lui; c.at(A.ru[1], x, y) = bu.x0; c.at(A.ru[2], x, y) = bu.y0; c.ati(A.ru[3], x, y) = pack_corner(bu.cx, bu.cy);
    c.at(A.rv[0], x, y) = lvi; c.at(A.rv[1], x, y) = bv.x0; c.at(A.rv[2], x, y) = bv.y0; c.ati(A.rv[3], x, y) = pack_corner(bv.cx, bv.cy);
    auto dist = [&](int cx) {
        int d = cx - x;
        if (PER) {                       // the wrapped distance
            d = ((d % nx) + nx) % nx;
            if (d > nx / 2) d -= nx;
        }
        return d < 0 ? -d : d;
    };
    int dx = amax(amax(dist(bs.cx), dist(bs.cx + 1)), amax(dist(bu.cx), dist(bu.cx + 1)));
    dx = amax(dx, amax(dist(bv.cx), dist(bv.cx + 1)));
    auto ady = [&](int cy) { return cy - y < 0 ? y - cy : cy - y; };
    int dy = amax(amax(ady(bs.cy), ady(bs.cy + 1)), amax(ady(bu.cy), ady(bu.cy + 1)));
    dy = amax(dy, amax(ady(bv.cy), ady(bv.cy + 1)));
    return dx | (dy << 16);
}

// Σ over the sources within (rx, ry) of (X, Y) of λ·(bilinear weight of corner (X, Y))
template <class R, bool PER, class C>
SWM_HD R tracer_gather(const C &c, const uint32_t rec[4], int X, int Y, int rx, int ry, int nx, int ny) {
    const R one = R(1);
    R acc = R(0);
    for (int dy = -ry; dy <= ry; ++dy) {
        const int y = Y - dy;
        if (y < 1 || y > ny) continue;
        for (int dx = -rx; dx <= rx; ++dx) {
            int x = X - dx;
            if (PER) x = ((x - 1) % nx + nx) % nx + 1;
            else if (x < 1 || x > nx) continue;
            const int p = c.ati(rec[3], x, y);
            const int cx = corner_x(p), cy = corner_y(p);
            const bool ix0 = X == cx, ix1 = X == cx + 1, iy0 = Y == cy, iy1 = Y == cy + 1;
            if (!(ix0 || ix1) || !(iy0 || iy1)) continue;
            const R x0 = c.at(rec[1], x, y), y0 = c.at(rec[2], x, y);
            const R wx = ix0 ? one - x0 : x0, wy = iy0 ? one - y0 : y0;
            acc = acc + c.at(rec[0], x, y) * (wx * wy);
        }
    }
    return acc;
}

// Second pass, in two parts.  (a) one thread per corner (X, Y), X = -2 .. nx+3,
// Y = -1 .. ny+3: λ of the old tracer, and the records of the um_T and vm_T interpolations
// gathered at the corner (kept in gu, gv: each is read at two corners by (b)).
template <class R, bool PER, class C>
SWM_HD void tracer_adj_gather_uv(const C &c, const TracerAdjArgs<R> &A, int X, int Y, int rx, int ry) {
    const int nx = A.nx, ny = A.ny;
    if (X >= 0 && X <= nx + 1 && Y >= 1 && Y <= ny)
        c.at(A.lsst_out, X, Y) = tracer_gather<R, PER>(c, A.rs, X, Y, rx, ry, nx, ny);
    c.at(A.gu, X, Y) = tracer_gather<R, PER>(c, A.ru, X, Y, rx + 1, ry + 1, nx, ny);
    c.at(A.gv, X, Y) = tracer_gather<R, PER>(c, A.rv, X, Y, rx + 1, ry + 1, nx, ny);
}

// (b) one thread per cell (X, Y) of the u/v/η/sst layouts (X = -2 .. nx+2, Y = -1 .. ny+2):
// accumulation into λu0, λv0 (through u_T, v_T) and λum, λvm (through um_T, vm_T)
template <class R, bool PER, class C>
SWM_HD void tracer_adj_gather(const C &c, const TracerAdjArgs<R> &A, int X, int Y) {
    const int nx = A.nx, ny = A.ny;
    const R half = R(0.5);
    const R lumT0 = c.at(A.gu, X, Y), lumT1 = c.at(A.gu, X + 1, Y);
    c.at(A.lum, X, Y) = c.at(A.lum, X, Y) + half * (lumT0 + lumT1);
    const R lvmT0 = c.at(A.gv, X, Y), lvmT1 = c.at(A.gv, X, Y + 1);
    c.at(A.lvm, X, Y) = c.at(A.lvm, X, Y) + half * (lvmT0 + lvmT1);
    auto inT = [&](int x, int y) { return x >= 1 && x <= nx && y >= 1 && y <= ny; };
    const R a0 = inT(X, Y) ? c.at(A.luT, X, Y) : R(0), a1 = inT(X + 1, Y) ? c.at(A.luT, X + 1, Y) : R(0);
    c.at(A.lu, X, Y) = c.at(A.lu, X, Y) + half * (a0 + a1);
    const R b0 = inT(X, Y) ? c.at(A.lvT, X, Y) : R(0), b1 = inT(X, Y + 1) ? c.at(A.lvT, X, Y + 1) : R(0);
    c.at(A.lv, X, Y) = c.at(A.lv, X, Y) + half * (b0 + b1);
}

// ---- the mid-point snapshot um ← u0 (tracer_advection.jl:12-17), adjoint: λu0 += λum,
// λum ← 0 (every cell of the layout, element k of the allocation)
template <class R> SWM_HD void snapshot_adjoint_elem(R *lu, R *lum, long k) {
    lu[k] = lu[k] + lum[k];
    lum[k] = R(0);
}

// ---- d = ax·x + ay·y + az·z at element k of the layout (every cell, halo included); y and z
// are left out when null.  d may be x.
template <class R> SWM_HD void lincomb_elem(R *d, const R *x, R ax, const R *y, R ay, const R *z, R az, long k) {
    R s = ax * x[k];
    if (y) s = s + ay * y[k];
    if (z) s = s + az * z[k];
    d[k] = s;
}

// ---- adjoint of the step's input: the ghost copies of the state folded into the interior
// and the halo cleared.  One thread per cell (x, y), x = -2 .. nx+2, y = -1 .. ny+2.
struct FoldOutArgs {
    UVE in, out;
    uint32_t sst_in, sst_out;
    int nx, ny, per;
};
template <class R, class C> SWM_HD void fold_out_cell(const C &c, const FoldOutArgs &A, R oma, int x, int y) {
    const int nx = A.nx, ny = A.ny;
    const bool per = A.per != 0;
    // u
    {
        const bool in = (per ? (x >= 0 && x <= nx - 1) : (x >= 1 && x <= nx - 1)) && y >= 1 && y <= ny;
        R l = R(0);
        if (in) {
            l = c.at(A.in.u, x, y);
            if (per) {
                if (x >= nx - 2) l = l + c.at(A.in.u, x - nx, y);
                if (x <= 1) l = l + c.at(A.in.u, x + nx, y);
            } else {
                if (y == 1) l = l + oma * c.at(A.in.u, x, 0);
                if (y == 2) l = l + oma * c.at(A.in.u, x, -1);
                if (y == ny) l = l + oma * c.at(A.in.u, x, ny + 1);
                if (y == ny - 1) l = l + oma * c.at(A.in.u, x, ny + 2);
            }
        }
        c.at(A.out.u, x, y) = l;
    }
    // v
    {
        const bool in = x >= 1 && x <= nx && y >= 1 && y <= ny - 1;
        R l = R(0);
        if (in) {
            l = c.at(A.in.v, x, y);
            if (per) {
                if (x >= nx - 1) l = l + c.at(A.in.v, x - nx, y);
                if (x <= 2) l = l + c.at(A.in.v, x + nx, y);
            } else {
                const int src = x == 1 ? 0 : x == 2 ? -1 : x == nx ? nx + 1 : x == nx - 1 ? nx + 2 : -100;
                if (src != -100) l = l + oma * c.at(A.in.v, src, y);
            }
        }
        c.at(A.out.v, x, y) = l;
    }
    // η
    {
        const bool in = x >= 1 && x <= nx && y >= 1 && y <= ny;
        R l = R(0);
        if (in) {
            const int xa = x == (per ? nx : 1) ? 0 : x == (per ? 1 : nx) ? nx + 1 : -100;
            const int ya = y == 1 ? 0 : y == ny ? ny + 1 : -100;
            l = c.at(A.in.e, x, y);
            if (xa != -100) l = l + c.at(A.in.e, xa, y);
            if (ya != -100) {
                R t = c.at(A.in.e, x, ya);
                if (xa != -100) t = t + c.at(A.in.e, xa, ya);
                l = l + t;
            }
        }
        c.at(A.out.e, x, y) = l;
    }
    // sst (ghost columns 0 and nx+1; no y halo)
    if (y >= 1 && y <= ny) {
        R l = R(0);
        if (x >= 1 && x <= nx) {
            l = c.at(A.sst_in, x, y);
            if (x == (per ? nx : 1)) l = l + c.at(A.sst_in, 0, y);
            if (x == (per ? 1 : nx)) l = l + c.at(A.sst_in, nx + 1, y);
        }
        c.at(A.sst_out, x, y) = l;
    }
}

// ------------------------------------------------------------------ the reverse of one step
//
// Buffers of the adjoint.  The tape holds the forward intermediates of the step being
// reversed (the stage inputs s_1..s_RKo, or under SSPRK2 / 4SSPRK3 the stage inputs w_k and
// the velocities after each stage's momentum update; u0 before and after the dissipation, the
// tracer state it read); L is the carried adjoint (halo unfolded), Z a second set of the same,
// Ma / Mb / A the μ of the classic RK stages RKo, RKo-1, RKo-2 (the SSP stages use Ma, Mb
// for the two marches of a stage and A for the kept state's adjoint).  The tracer adjoint's
// records reuse Z, Ma, Mb, A and two more fields.
struct AdjFields {
    UVE s[12];                      // stage inputs (classic RK: 4 at most; SSPRK3 with n = 3: 10)
    UVE p[12];                      // SSP schemes: the velocities of stage k after its momentum update (u, v)
    UVE pre;
    uint32_t post_u, post_v, um, vm, sst;
    UVE L, Z, Ma, Mb, A;
    uint32_t Lsst, Lsst2, Lum, Lvm;
    uint32_t H, fq;
    int fq_2d;
    uint32_t fu, fv;                // linear dynamics: f on the u- and v-grid
    int fu_2d, fv_2d;
    uint32_t qk, lqk, lkk;          // kept q the step's stages read (taped), λ of the kept q and KEu+KEv
    uint32_t qend, lqend;
    uint32_t tr0, tr1;
    uint32_t tg0, tg1;              // the tracer adjoint's gathered interpolation records
    uint32_t lfx;                   // λ of Fx at the u points (accumulated over the reversed steps)
    uint32_t zero;                  // a field of zeros (never written)
};

// Scalars of the configuration the reverse step needs (all in the arithmetic type R).
template <class R> struct AdjConsts {
    R a[4], b[4];                   // RKaΔt, RKbΔt (b[3] unused)
    R g, gamma, k12, oma;
    R cD, rD, cS, nuB, dtd;
    R hdu, hdv, du, dv, jSST;
    int RKo, cont, drag, smag, per, consume, tracer, nadvstep, nadvstep_half, nstep_diff;
    int nx, ny, P, ep;
    R fxt;                          // the step's factor of Fx in du (1 unless the wind is seasonal)
    int fxg;                        // accumulate the cotangent of Fx
    int sad;                        // Sadourny's PV fluxes (k12 = ½)
    int lin;                        // the linear dynamics (rhs_linear!)
    int nadvcor;                    // nstep_advcor: q and KEu+KEv kept (0: formed by every stage)
    int ssp;                        // SSPRK2: the number of stages RKs (0: classic RK)
    int s4;                         // 4SSPRK3
    int s3, kn, mn;                 // SSPRK3: its s stages; the kn-th (counted from 1) combines with the
                                    // state after the mn-th (mn = 0: the step's input)
    R cs, sa, sb;                   // SSPRK2: Δt_Δs, and the weights 1/RKs, (RKs-1)/RKs of u0 = a·u + b·u1;
                                    // 4SSPRK3: ½·Δt_Δ, and the weights 2/3, 1/3 of the third stage's combination;
                                    // SSPRK3: Δt_Δn, and kna, knb of the kn-th stage
    R cs2;                          // SSPRK3: Δt_Δnc, the kn-th stage's factor of its tendency
};

// The reverse of step i: L (the adjoint of the state after the step) becomes the adjoint of
// the state before it.  Launcher K issues the kernels: stage(args, per, next, nx, pre, form),
// qend(args), qend_adj(args, mu), diss<PER, DRAG, SMAG>(args), tracer(args), snapshot(),
// advcor_adj(args), lqend_keep(args), zero(field), lincomb(d, x, ax, y, ay, z, az), in order
// on one queue.
template <class R, class K> int reverse_tracer(K &k, AdjFields &F, const AdjConsts<R> &c, long i) {
    int rc = 0;
    // tracer! (tracer_advection.jl:1-44): advection (:19-42), then the snapshot (:12-17)
    if (c.tracer) {
        if (i % c.nadvstep == 0) {
            TracerAdjArgs<R> T{};
            T.u = F.post_u; T.v = F.post_v; T.um = F.um; T.vm = F.vm; T.sst = F.sst;
            T.lsst = F.Lsst; T.lsst_out = F.Lsst2;
            T.lu = F.L.u; T.lv = F.L.v; T.lum = F.Lum; T.lvm = F.Lvm;
            const uint32_t rec[12] = {F.Z.u, F.Z.v, F.Z.e, F.Ma.u, F.Ma.v, F.Ma.e, F.Mb.u, F.Mb.v, F.Mb.e, F.A.u, F.A.v, F.A.e};
            for (int j = 0; j < 4; ++j) { T.rs[j] = rec[j]; T.ru[j] = rec[4 + j]; T.rv[j] = rec[8 + j]; }
            T.luT = F.tr0; T.lvT = F.tr1; T.gu = F.tg0; T.gv = F.tg1;
            T.hdu = c.hdu; T.hdv = c.hdv; T.du = c.du; T.dv = c.dv; T.jSST = c.jSST;
            T.consume = c.consume; T.nx = c.nx; T.ny = c.ny; T.P = c.P; T.ep = c.ep;
            if ((rc = k.tracer(T))) return rc;
            const uint32_t t = F.Lsst; F.Lsst = F.Lsst2; F.Lsst2 = t;
        }
        if ((i + c.nadvstep_half) % c.nadvstep == 0 && (rc = k.snapshot(F.L.u, F.Lum, F.L.v, F.Lvm))) return rc;
    }
    return rc;
}

constexpr uint32_t NOF = 0xFFFFFFFFu;   // "no field" (lincomb operands left out)

// The stages of an SSPRK2 or 4SSPRK3 step, last first.  Both move the velocities first and
// then η with the mass fluxes of the updated velocities (the semi-implicit continuity):
//   SSPRK2 (time_integration.jl:80-110), stage k from w_k = (u_k, v_k, η_k):
//     u'_k = u_k + c·du(w_k), v likewise, ghost_points_uv!;  η'_k = η_k + c·dη(u'_k, v'_k, η_k)
//     w_{k+1} = (u'_k, v'_k, η'_k) (η ghost copies at the next stage); u0 = a·u + b·w_{RKs+1}
//     (c = Δt_Δs, a = 1/RKs, b = (RKs-1)/RKs).
//   4SSPRK3 (:158-195): u'_k = ½(u_k + (u_k + Δt_Δ·du(w_k))), η'_k likewise, i.e. weight 1 on the
//     input and c = ½·Δt_Δ on the tendency; w_{k+1} = (u'_k, v'_k, η'_k) except
//     w_4 = ⅔·w_1 + ⅓·(u'_3, v'_3, η'_3); u0 = (u'_4, v'_4, η'_4).
//   SSPRK3 (:112-156), s stages: weight 1 and c = Δt_Δn, except the kn-th stage, which has
//     weight kna on its input, c = Δt_Δnc, and adds knb times the state X kept after the mn-th
//     stage (the step's input when mn = 0); w_{k+1} = (u'_k, v'_k, η'_k); u0 = w_{s+1}.
// ghost_points! on u0 follows.  Every adjoint is carried with its halo unfolded (a halo cell
// holds the adjoint of the ghost copy, folded into the cell it copies by its reader), so the
// stage bookkeeping is elementwise.  With Λ the adjoint of the stage variables and λ0 that of u0
// (l0), and (id_k, c_k) a stage's weights:
//   Λ ← b·λ0 (SSPRK2) or λ0;  for k = last .. 1:
//     (4SSPRK3, k = 3: A ← ⅔·Λ, Λ ← ⅓·Λ)
//     μ2 = (∂dη/∂(u'_k, v'_k, η_k))ᵀ c_k·Λη                (stage march, continuity form)
//     Λu += μ2u, Λv += μ2v                                 (Λ of u'_k, v'_k complete)
//     (SSPRK3, k = kn: B ← knb·Λ, the adjoint of X)
//     μ1 = (∂(du, dv)/∂w_k)ᵀ c_k·(Λu, Λv)                  (stage march with λη = 0)
//     Λu ← id_k·Λu + μ1u, Λv likewise, Λη ← id_k·Λη + μ2η + μ1η   (Λ of w_k)
//     (SSPRK3, k = mn + 1: Λ += B, as w_k is X)
//   the step's input: Λ + a·λ0 (SSPRK2), Λ + A (4SSPRK3) or Λ (SSPRK3).
// The taped fields: F.s[k-1] = w_k, F.p[k-1] = (u'_k, v'_k).  The result goes to `out`; F.L
// and F.Z are swapped as by the classic RK stages.
template <class R, class K> int reverse_ssp(K &k, AdjFields &F, const AdjConsts<R> &c, UVE l0, UVE out) {
    int rc = 0;
    const bool s4 = c.s4 != 0, s3 = c.s3 > 0;
    const int ns = s4 ? 4 : s3 ? c.s3 : c.ssp;
    const UVE Lm = out, M2 = F.Ma, M1 = F.Mb;
    const R one = R(1), nil = R(0), b = c.ssp > 0 ? c.sb : one;
    if ((rc = k.lincomb(Lm.u, l0.u, b, NOF, nil, NOF, nil)) || (rc = k.lincomb(Lm.v, l0.v, b, NOF, nil, NOF, nil)) ||
        (rc = k.lincomb(Lm.e, l0.e, b, NOF, nil, NOF, nil)))
        return rc;
    StageAdjArgs<R> S{};
    S.H = F.H; S.fq = F.fq; S.fq_2d = F.fq_2d; S.qend = F.qend; S.lqend = F.lqend;
    S.fu = F.fu; S.fv = F.fv; S.fu_2d = F.fu_2d; S.fv_2d = F.fv_2d;
    S.qk = F.qk; S.lqk = F.lqk; S.lkk = F.lkk;
    S.g = c.g; S.gamma = c.gamma; S.k12 = c.k12; S.oma = c.oma; S.cont = c.cont;
    S.nx = c.nx; S.ny = c.ny; S.P = c.P;
    S.lfx = F.lfx; S.fxt = c.fxt;
    S.b = nil;
    S.mu_in = S.a_out = S.mu_x[0] = S.mu_x[1] = F.A;   // (read and written by NEXT stages only)
    const int sch = (c.lin ? ADJ_LIN : c.sad ? ADJ_SAD : ADJ_AH) | (c.nadvcor > 0 && !c.lin ? ADJ_KEEP : 0);
    const bool wrap = c.per && (sch & 3) == ADJ_AH, keep = (sch & ADJ_KEEP) != 0;
    for (int st = ns; st >= 1; --st) {
        const bool kn = s3 && st == c.kn;
        const R id = kn ? c.sa : one;
        S.a = kn ? c.cs2 : c.cs;
        if (s4 && st == 3) {                        // w_4 = ⅔·w_1 + ⅓·(u'_3, v'_3, η'_3)
            if ((rc = k.lincomb(F.A.u, Lm.u, c.sa, NOF, nil, NOF, nil)) || (rc = k.lincomb(F.A.v, Lm.v, c.sa, NOF, nil, NOF, nil)) ||
                (rc = k.lincomb(F.A.e, Lm.e, c.sa, NOF, nil, NOF, nil)) || (rc = k.lincomb(Lm.u, Lm.u, c.sb, NOF, nil, NOF, nil)) ||
                (rc = k.lincomb(Lm.v, Lm.v, c.sb, NOF, nil, NOF, nil)) || (rc = k.lincomb(Lm.e, Lm.e, c.sb, NOF, nil, NOF, nil)))
                return rc;
        }
        // continuity at (u'_k, v'_k, η_k)
        S.s = UVE{F.p[st - 1].u, F.p[st - 1].v, F.s[st - 1].e};
        S.l0 = Lm; S.mu_out = M2; S.fxg = 0;
        if ((rc = k.stage(S, c.per, 0, 0, 0, ADJ_CONT))) return rc;
        if ((rc = k.lincomb(Lm.u, Lm.u, one, M2.u, one, NOF, nil)) || (rc = k.lincomb(Lm.v, Lm.v, one, M2.v, one, NOF, nil)))
            return rc;
        if (kn && ((rc = k.lincomb(F.A.u, Lm.u, c.sb, NOF, nil, NOF, nil)) ||
                   (rc = k.lincomb(F.A.v, Lm.v, c.sb, NOF, nil, NOF, nil)) || (rc = k.lincomb(F.A.e, Lm.e, c.sb, NOF, nil, NOF, nil))))
            return rc;
        // momentum at w_k
        S.s = F.s[st - 1];
        S.l0 = UVE{Lm.u, Lm.v, F.zero}; S.mu_out = M1; S.fxg = c.fxg;
        if (wrap && !keep && (rc = k.qend(S))) return rc;
        if ((rc = k.stage(S, c.per, 0, 0, 0, sch))) return rc;
        if (wrap && !keep && (rc = k.qend_adj(S, M1))) return rc;
        if (wrap && keep && (rc = k.lqend_keep(S))) return rc;
        if ((rc = k.lincomb(Lm.u, Lm.u, id, M1.u, one, NOF, nil)) || (rc = k.lincomb(Lm.v, Lm.v, id, M1.v, one, NOF, nil)) ||
            (rc = k.lincomb(Lm.e, Lm.e, id, M2.e, one, M1.e, one)))
            return rc;
        if (s3 && st == c.mn + 1 &&
            ((rc = k.lincomb(Lm.u, Lm.u, one, F.A.u, one, NOF, nil)) || (rc = k.lincomb(Lm.v, Lm.v, one, F.A.v, one, NOF, nil)) ||
             (rc = k.lincomb(Lm.e, Lm.e, one, F.A.e, one, NOF, nil))))
            return rc;
    }
    if (!s3) {
        const UVE in = s4 ? F.A : l0;
        const R w = s4 ? one : c.sa;
        if ((rc = k.lincomb(Lm.u, Lm.u, one, in.u, w, NOF, nil)) || (rc = k.lincomb(Lm.v, Lm.v, one, in.v, w, NOF, nil)) ||
            (rc = k.lincomb(Lm.e, Lm.e, one, in.e, w, NOF, nil)))
            return rc;
    }
    F.L = out;
    F.Z = l0;
    return 0;
}

template <class R, class K> int reverse_step(K &k, AdjFields &F, const AdjConsts<R> &c, long i) {
    int rc = reverse_tracer<R>(k, F, c, i);
    if (rc) return rc;
    // dissipation (time_integration.jl:216-221)
    UVE l0 = F.L, out = F.Z;
    const bool diss = i % c.nstep_diff == 0;
    if (diss) {
        DissAdjArgs<R> D{};
        D.s = F.pre; D.H = F.H; D.lin = F.L; D.out = F.Z;
        D.cD = c.cD; D.rD = c.rD; D.cS = c.cS; D.nuB = c.nuB; D.dtd = c.dtd; D.oma = c.oma;
        D.nx = c.nx; D.ny = c.ny; D.P = c.P;
        if ((rc = k.diss(D, c.per, c.drag, c.smag))) return rc;
        l0 = F.Z;
        out = F.L;
    }
    // the kept q and KEu+KEv formed from the state after the stages (time_integration.jl:208-211;
    // KEu+KEv again by the quadratic drag): their cotangents, gathered from the steps that read
    // them, pass to that state, and the stages of this step read the values formed before
    if (c.nadvcor > 0) {
        const bool wq = i % c.nadvcor == 0, wk = wq || (diss && c.drag == 0);
        if (wk) {
            AdvcorAdjArgs X{};
            X.s = F.pre; X.H = F.H; X.fq = F.fq; X.fq_2d = F.fq_2d; X.lqk = F.lqk; X.lkk = F.lkk; X.out = l0;
            X.wq = wq; X.nx = c.nx; X.ny = c.ny;
            if ((rc = k.advcor_adj(X))) return rc;
            if ((rc = k.zero(F.lkk))) return rc;
            if (wq && (rc = k.zero(F.lqk))) return rc;
        }
    }
    if (c.ssp > 0 || c.s4 || c.s3 > 0) return reverse_ssp(k, F, c, l0, out);
    // the RK stages, last first (time_integration.jl:55-78)
    StageAdjArgs<R> S{};
    S.H = F.H; S.fq = F.fq; S.fq_2d = F.fq_2d; S.qend = F.qend; S.lqend = F.lqend;
    S.fu = F.fu; S.fv = F.fv; S.fu_2d = F.fu_2d; S.fv_2d = F.fv_2d;
    S.qk = F.qk; S.lqk = F.lqk; S.lkk = F.lkk;
    const int sch = (c.lin ? ADJ_LIN : c.sad ? ADJ_SAD : ADJ_AH) | (c.nadvcor > 0 && !c.lin ? ADJ_KEEP : 0);
    S.l0 = l0;
    S.g = c.g; S.gamma = c.gamma; S.k12 = c.k12; S.oma = c.oma; S.cont = c.cont;
    S.nx = c.nx; S.ny = c.ny; S.P = c.P;
    S.lfx = F.lfx; S.fxt = c.fxt; S.fxg = c.fxg;
    // μ of stages RKo, RKo-1, RKo-2 in three buffers: the first stage reads them all
    const UVE mu[3] = {F.Ma, F.Mb, F.A};
    S.mu_x[0] = mu[0];
    S.mu_x[1] = mu[1];
    S.a_out = out;
    for (int st = c.RKo; st >= 1; --st) {
        const bool next = st < c.RKo, pre = st == 1;
        const int nx = pre ? c.RKo - 2 : 0;
        S.s = F.s[st - 1];
        S.a = c.a[st - 1];
        S.b = next ? c.b[st - 1] : R(0);
        S.mu_in = next ? mu[c.RKo - st - 1] : F.A;
        S.mu_out = pre ? out : mu[c.RKo - st];
        const bool wrap = c.per && (sch & 3) == ADJ_AH; // the Arakawa-Hsu first-column wrap to q[end,·]
        const bool keep = (sch & ADJ_KEEP) != 0;        // (kept q: q[end,·] is the kept field's)
        if (wrap && !keep && (rc = k.qend(S))) return rc;
        if ((rc = k.stage(S, c.per, next, nx, pre, sch))) return rc;
        if (wrap && !keep && (rc = k.qend_adj(S, S.mu_out))) return rc;
        if (wrap && keep && (rc = k.lqend_keep(S))) return rc;
    }
    F.L = out;
    F.Z = l0;
    return 0;
}

}  // namespace adj
}  // namespace swm
