// swm_common.h -- device layout, cross-lane moves, ghost copies and wave tiling shared by
// the gfx950 kernels of the ShallowWaters.jl time step.
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
// barriers); every stencil of the step has an x-radius of at most 2, so lanes 2..61 of a
// strip produce output and strips advance by 60 columns.  Each output is written once,
// together with the ghost copies the reference's ghost_points! makes from it.
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

// ------------------------------------------------------------------ static fields

template <class R> struct Static {
    const R *H;                       // T-point, 2-D
    const R *Fx; int Fx_sx, Fx_sy;    // U-point; sx=0 when x-invariant (1-D in y)
    const R *Fy; int Fy_sx, Fy_sy;    // V-point
    const R *fq; int fq_sx, fq_sy;    // Q-point
    const R *eta_ref, *Feta;          // T-point, interior (continuity relax / forcing)
};

// continuity variant (continuity.jl:91-97), a runtime switch uniform over the launch
enum { CONT_ITSELF = 0, CONT_RELAX = 1, CONT_FORCING = 2 };

// dη at T-point (x, y) from the flux divergence, continuity.jl:3-68
template <class R>
__device__ __forceinline__ R continuity_rhs(R dUdx, R dVdy, R e, int cont, const Static<R> &st, int x, int y, int P,
                                            R gamma, R Fet) {
    R d = -(dUdx + dVdy);
    if (cont == CONT_RELAX) d = d + gamma * (st.eta_ref[x + y * P] - e);
    else if (cont == CONT_FORCING) d = d + Fet * st.Feta[x + y * P];
    return d;
}

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

}  // namespace swm
