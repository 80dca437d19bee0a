// swm_common.h -- device layout, field access, cross-lane moves, ghost copies and wave
// tiling shared by the gfx950 kernels of the ShallowWaters.jl time step.
//
// Device layout.  Every field, whatever its staggering, lives in one padded 2-D layout
// indexed by T-cell coordinates (x, y): element (x, y) sits at element offset
// (x + OX) + (y + OY)*P of its allocation.  A field's staggering is implicit:
//   T-point  (η, h, H, p, ...)  centre of cell (x, y)
//   U-point  (u, U, du, Fx ...) right face of cell (x, y)
//   V-point  (v, V, dv, Fy ...) top face of cell (x, y)
//   Q-point  (q, f_q, ...)      top-right corner of cell (x, y)
// so every reference stencil becomes a fixed offset pattern in {-1,0,1}^2 and the
// reference's ep-dependent index shifts disappear into the mapping Julia index <-> (x, y):
//   u[ii,jj] -> (ii-2-ep, jj-2)   v[ii,jj] -> (ii-2, jj-2)   η[ii,jj] -> (ii-1, jj-1)
//   q[ii,jj] -> (ii-1, jj-1)      sst[ii,jj] -> (ii-1, jj)
// The padding (x from -64, y from -4) keeps every lane's access in bounds; padding is zero
// and never written.
//
// Field access goes through buffer resources: a lane's byte offset of its column is a
// loop-invariant VGPR and the row is a scalar offset, so marching a row costs one scalar
// add instead of per-lane 64-bit address arithmetic.
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
constexpr int NXCD = 8;      // accelerator complex dies of an MI355X, each with its own L2

// ------------------------------------------------------------------ field access

// All fields of a context live in one device arena; a kernel addresses them through one
// buffer descriptor.  A field is named by the byte offset of its allocation in the arena
// (its origin), element (x, y) at origin + colb(x) + rowb(y): the column part is a
// loop-invariant VGPR, origin + row a scalar, so marching a row costs one scalar add.
struct ArenaRef {
    const void *base;
    uint32_t bytes;
};

template <class R> struct Arena {
    __amdgpu_buffer_rsrc_t rs;
    __device__ __forceinline__ explicit Arena(ArenaRef a)
        : rs(__builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(a.base), 0, (int)a.bytes, 0x00020000)) {}
    // vo: per-lane byte offset, so: uniform byte offset (field origin + row)
    __device__ __forceinline__ R ld(uint32_t vo, uint32_t so) const {
        if constexpr (sizeof(R) == 4) return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b32(rs, vo, so, 0));
        else return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b64(rs, vo, so, 0));
    }
    __device__ __forceinline__ void st(R v, uint32_t vo, uint32_t so) const {
        if constexpr (sizeof(R) == 4) __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), rs, vo, so, 0);
        else {
            using u2 = decltype(__builtin_amdgcn_raw_buffer_load_b64(rs, 0, 0, 0));
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), rs, vo, so, 0);
        }
    }
};

// a field of the arena: its origin and the arena it lives in
template <class R> struct Fld {
    const Arena<R> *a;
    uint32_t o;
    __device__ __forceinline__ R ld(uint32_t vo, uint32_t so) const { return a->ld(vo, o + so); }
    __device__ __forceinline__ void st(R v, uint32_t vo, uint32_t so) const { a->st(v, vo, o + so); }
};

// byte offsets of column x and row y in the padded layout
template <class R> __device__ __forceinline__ uint32_t colb(int x) { return (uint32_t)(x + OX) * (uint32_t)sizeof(R); }
template <class R> __device__ __forceinline__ uint32_t rowb(int y, int P) {
    return (uint32_t)(y + OY) * (uint32_t)P * (uint32_t)sizeof(R);
}

// ------------------------------------------------------------------ cross-lane moves
// value of lane+1 (x+1) and lane-1 (x-1); the edge lanes receive junk that only feeds
// the two halo lanes on each side.
// Whole-wave DPP shifts (a modifier on a VALU move, no LDS round trip):
// wave_shl:1 (0x130) gives lane i the value of lane i+1, wave_shr:1 (0x138) that of lane i-1.
constexpr int DPP_WAVE_SHL1 = 0x130, DPP_WAVE_SHR1 = 0x138;
template <int CTRL> __device__ __forceinline__ float dpp_f(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
}
template <int CTRL> __device__ __forceinline__ double dpp_f(double v) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, 0xF, 0xF, true);
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}
__device__ __forceinline__ float sh_xp(float v) { return dpp_f<DPP_WAVE_SHL1>(v); }
__device__ __forceinline__ float sh_xm(float v) { return dpp_f<DPP_WAVE_SHR1>(v); }
__device__ __forceinline__ double sh_xp(double v) { return dpp_f<DPP_WAVE_SHL1>(v); }
__device__ __forceinline__ double sh_xm(double v) { return dpp_f<DPP_WAVE_SHR1>(v); }

// ------------------------------------------------------------------ static fields

struct StaticRefs {
    uint32_t H;                        // T-point, 2-D
    uint32_t Fx, Fy, fq;               // U-, V-, Q-point; 1-D per row when x-invariant
    int Fx_2d, Fy_2d, fq_2d;           // 1: 2-D layout, 0: one value per row
    uint32_t eta_ref, Feta;            // T-point interior (continuity relax / forcing)
};

// byte offset of (x, y) in a static field that may be stored 1-D (one value per row)
template <class R> __device__ __forceinline__ uint32_t scolb(int x, int is2d) { return is2d ? colb<R>(x) : 0u; }
template <class R> __device__ __forceinline__ uint32_t srowb(int y, int P, int is2d) {
    return is2d ? rowb<R>(y, P) : (uint32_t)(y + OY) * (uint32_t)sizeof(R);
}

// continuity variant (continuity.jl:91-97), a runtime switch uniform over the launch
enum { CONT_ITSELF = 0, CONT_RELAX = 1, CONT_FORCING = 2 };

// ------------------------------------------------------------------ ghost copies on store
// ghost_points.jl:47-133: each interior value is written together with the halo cells the
// reference copies from it.  The common case (no ghost cell) costs one test of a flag
// precomputed per lane (x-copies) and a uniform row test (y-copies).

// lanes whose u, v or η value is the source of an x-direction ghost copy
struct GhostLanes {
    bool u, v, e;
};
template <bool PER> __device__ __forceinline__ GhostLanes ghost_lanes(int x, int nx) {
    GhostLanes g;
    g.u = PER && (x == 0 || x == 1 || x == nx - 2 || x == nx - 1);
    g.v = x == 1 || x == 2 || x == nx - 1 || x == nx;
    g.e = x == 1 || x == nx;
    return g;
}

template <class R, bool PER>
__device__ __forceinline__ void ghost_u(const Fld<R> &b, int x, int y, R w, int nx, int ny, int P, R oma) {
    if (PER) {                                   // ghost_points_u_periodic!  :60-73
        const uint32_t so = rowb<R>(y, P);
        if (x == nx - 2) b.st(w, colb<R>(-2), so);
        if (x == nx - 1) b.st(w, colb<R>(-1), so);
        if (x == 0) b.st(w, colb<R>(nx), so);
        if (x == 1) b.st(w, colb<R>(nx + 1), so);
    } else {                                     // ghost_points_u_nonperiodic!  :47-57
        const R gw = oma * w;
        const uint32_t vo = colb<R>(x);
        if (y == 1) b.st(gw, vo, rowb<R>(0, P));
        if (y == 2) b.st(gw, vo, rowb<R>(-1, P));
        if (y == ny) b.st(gw, vo, rowb<R>(ny + 1, P));
        if (y == ny - 1) b.st(gw, vo, rowb<R>(ny + 2, P));
    }
}

template <class R, bool PER>
__device__ __forceinline__ void ghost_v(const Fld<R> &b, int x, int y, R w, int nx, int P, R oma) {
    const uint32_t so = rowb<R>(y, P);
    if (PER) {                                   // ghost_points_v_periodic!  :89-99
        if (x == nx - 1) b.st(w, colb<R>(-1), so);
        if (x == nx) b.st(w, colb<R>(0), so);
        if (x == 1) b.st(w, colb<R>(nx + 1), so);
        if (x == 2) b.st(w, colb<R>(nx + 2), so);
    } else {                                     // ghost_points_v_nonperiodic!  :76-86
        const R gw = oma * w;
        if (x == 1) b.st(gw, colb<R>(0), so);
        if (x == 2) b.st(gw, colb<R>(-1), so);
        if (x == nx) b.st(gw, colb<R>(nx + 1), so);
        if (x == nx - 1) b.st(gw, colb<R>(nx + 2), so);
    }
}

template <class R, bool PER>
__device__ __forceinline__ void ghost_e(const Fld<R> &b, int x, int y, R w, int nx, int ny, int P) {
    // ghost_points_η_(non)periodic!  :102-133: x-copies first, then y-copies of whole rows
    const bool gy0 = y == 1, gy1 = y == ny;
    const uint32_t s0 = rowb<R>(0, P), s1 = rowb<R>(ny + 1, P), so = rowb<R>(y, P);
    const uint32_t vo = colb<R>(x);
    if (gy0) b.st(w, vo, s0);
    if (gy1) b.st(w, vo, s1);
    if (x == (PER ? nx : 1)) {
        const uint32_t vg = colb<R>(0);
        b.st(w, vg, so);
        if (gy0) b.st(w, vg, s0);
        if (gy1) b.st(w, vg, s1);
    }
    if (x == (PER ? 1 : nx)) {
        const uint32_t vg = colb<R>(nx + 1);
        b.st(w, vg, so);
        if (gy0) b.st(w, vg, s0);
        if (gy1) b.st(w, vg, s1);
    }
}

// store of an interior value plus its ghost copies
template <class R, bool PER>
__device__ __forceinline__ void put_u(const Fld<R> &b, int x, uint32_t vo, int y, uint32_t so, R w, bool glane, int nx,
                                      int ny, int P, R oma) {
    b.st(w, vo, so);
    if (PER ? glane : (y <= 2 || y >= ny - 1)) ghost_u<R, PER>(b, x, y, w, nx, ny, P, oma);
}
template <class R, bool PER>
__device__ __forceinline__ void put_v(const Fld<R> &b, int x, uint32_t vo, int y, uint32_t so, R w, bool glane, int nx,
                                      int P, R oma) {
    b.st(w, vo, so);
    if (glane) ghost_v<R, PER>(b, x, y, w, nx, P, oma);
}
template <class R, bool PER>
__device__ __forceinline__ void put_e(const Fld<R> &b, int x, uint32_t vo, int y, uint32_t so, R w, bool glane, int nx,
                                      int ny, int P) {
    b.st(w, vo, so);
    if (glane || y == 1 || y == ny) ghost_e<R, PER>(b, x, y, w, nx, ny, P);
}

// ------------------------------------------------------------------ wave -> (strip, segment)

struct WaveTile {
    int lane, x, ya, yb;
    bool live;
};

__device__ __forceinline__ WaveTile wave_tile(int nstrips, int nseg, int L, int ny) {
    WaveTile t;
    // Blocks are dispatched round-robin over the XCDs (block b to XCD b % NXCD).  The grid is
    // a multiple of NXCD blocks; renumbering gives each XCD a contiguous run of tiles, so the
    // halo rows two segments share sit in the same L2.
    const int per = (int)gridDim.x / NXCD;
    const int b = ((int)blockIdx.x % NXCD) * per + (int)blockIdx.x / NXCD;
    // the wave index is uniform over the wave: readfirstlane makes it (and the rows) scalar
    const int w = b * WAVES_PER_BLOCK + __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    t.lane = (int)(threadIdx.x % WAVE);
    t.live = w < nstrips * nseg;
    const int s = w % nstrips, g = w / nstrips;
    t.x = -HALO + WOUT * s + t.lane;
    t.ya = 1 + g * L;
    t.yb = min(t.ya + L, ny + 1);
    return t;
}

}  // namespace swm
