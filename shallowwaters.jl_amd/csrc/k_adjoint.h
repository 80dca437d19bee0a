// k_adjoint.h -- gfx950 kernels of the adjoint time step: the marches and per-cell passes of
// adj_core.h on one wavefront per strip (lanes are columns, DPP shifts for x±1, scalar row
// offsets), and the small per-row / per-cell kernels around them.
#pragma once

#include "adj_core.h"
#include "swm_common.h"

namespace swm {

static_assert(adj::AOX == OX && adj::AOY == OY && adj::AWAVE == WAVE, "the adjoint uses the field layout of the time step");

// One wavefront of a march: lane l holds column x_lane0 + l.
template <class R_> struct GpuWave {
    using R = R_;
    using V = R;
    using M = bool;
    using I = int;
    Arena ar;
    int P;
    I x;
    M own;
    uint32_t vx;
    __device__ __forceinline__ GpuWave(ArenaRef a, int P_, int x_lane0, int H) : ar(a), P(P_) {
        const int lane = (int)(threadIdx.x % WAVE);
        x = x_lane0 + lane;
        own = lane >= H && lane < WAVE - H;
        vx = colb<R>(x);
    }
    __device__ __forceinline__ V ld(uint32_t o, int y) const { return Sto<R>::ld(ar, vx, o + rowb<R>(y, P)); }
    __device__ __forceinline__ V ld_at(uint32_t o, I xc, int y, M m) const {
        return m ? Sto<R>::ld(ar, colb<R>(xc), o + rowb<R>(y, P)) : R(0);
    }
    __device__ __forceinline__ V ld_row1(uint32_t o, int y) const {
        return Sto<R>::ld(ar, 0, o + (uint32_t)(y + OY) * (uint32_t)sizeof(R));
    }
    __device__ __forceinline__ V ld_s(uint32_t o, int y, int is2d) const { return is2d ? ld(o, y) : ld_row1(o, y); }
    __device__ __forceinline__ void st(uint32_t o, int y, V v, M m) const {
        if (m) Sto<R>::st(ar, v, vx, o + rowb<R>(y, P));
    }
    __device__ __forceinline__ void st_row1(uint32_t o, int y, V v, M m) const {
        if (m) Sto<R>::st(ar, v, 0, o + (uint32_t)(y + OY) * (uint32_t)sizeof(R));
    }
    __device__ __forceinline__ static V xp(V v) { return sh_xp(v); }
    __device__ __forceinline__ static V xm(V v) { return sh_xm(v); }
    __device__ __forceinline__ static V sel(M m, V a, V b) { return m ? a : b; }
    __device__ __forceinline__ static I seli(M m, int a, I b) { return m ? a : b; }
    __device__ __forceinline__ static V vsqrt(V v) { return nsqrt(v); }
};

// cell access through the arena's base pointer (per-cell and per-row passes)
template <class R> struct GpuCells {
    char *base;
    int P;
    __device__ __forceinline__ R &at(uint32_t o, int x, int y) const {
        return *(R *)(base + o + ((size_t)(x + OX) + (size_t)(y + OY) * (size_t)P) * sizeof(R));
    }
    __device__ __forceinline__ int &ati(uint32_t o, int x, int y) const {
        return *(int *)(base + o + ((size_t)(x + OX) + (size_t)(y + OY) * (size_t)P) * sizeof(R));
    }
    __device__ __forceinline__ R &row1(uint32_t o, int y) const { return *(R *)(base + o + (size_t)(y + OY) * sizeof(R)); }
};

template <class R, bool PER, bool NEXT, int NX, bool PRE, int SCH = adj::ADJ_AH>
__global__ void __launch_bounds__(256) stage_adj_kernel(ArenaRef ar, adj::StageAdjArgs<R> A) {
    const adj::Tile t = adj::tile_of<adj::RT_HALO>(wave_index(), A.nstrips, A.nseg, A.L, adj::ADJ_Y0, adj::adj_y1(A.ny));
    if (!t.live) return;
    GpuWave<R> w(ar, A.P, t.x_lane0, adj::RT_HALO);
    adj::stage_adjoint<GpuWave<R>, PER, NEXT, NX, PRE, SCH>(w, A, t.ya, t.yb);
}

template <class R, bool PER, int DRAG, bool SMAG>
__global__ void __launch_bounds__(256) diss_adj_kernel(ArenaRef ar, adj::DissAdjArgs<R> A) {
    const adj::Tile t = adj::tile_of<adj::DT_HALO>(wave_index(), A.nstrips, A.nseg, A.L, adj::ADJ_Y0, adj::adj_y1(A.ny));
    if (!t.live) return;
    GpuWave<R> w(ar, A.P, t.x_lane0, adj::DT_HALO);
    adj::diss_adjoint<GpuWave<R>, PER, DRAG, SMAG>(w, A, t.ya, t.yb);
}

template <class R> __global__ void qend_fwd_kernel(ArenaRef ar, adj::StageAdjArgs<R> A) {
    const int y = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (y > A.ny) return;
    adj::qend_row<R>(GpuCells<R>{(char *)ar.base, A.P}, A, y);
}

// the kept q and KEu+KEv (nstep_advcor > 0): their cotangents into the state's, one thread
// per cell x in -2 .. nx+2, y in -1 .. ny+2
template <class R> __global__ void advcor_adj_kernel(ArenaRef ar, adj::AdvcorAdjArgs A, int P) {
    const int x = -2 + (int)(blockIdx.x * blockDim.x + threadIdx.x), y = -1 + (int)blockIdx.y;
    if (x > A.nx + 2) return;
    adj::advcor_adjoint_cell<R>(GpuCells<R>{(char *)ar.base, P}, A, x, y);
}

template <class R> __global__ void lqend_keep_kernel(ArenaRef ar, adj::StageAdjArgs<R> A) {
    const int y = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (y > A.ny) return;
    adj::lqend_keep_row<R>(GpuCells<R>{(char *)ar.base, A.P}, A, y);
}

template <class R> __global__ void qend_adj_kernel(ArenaRef ar, adj::StageAdjArgs<R> A, adj::UVE mu) {
    const int yy = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (yy > A.ny + 1) return;
    adj::qend_adjoint_row<R>(GpuCells<R>{(char *)ar.base, A.P}, A, mu, yy);
}

// first tracer pass, one thread per T-column over `rows` rows; dmax[0] (dmax[1]) receives the
// largest x (y) distance between a cell and the corners it interpolates from.  Atomics on
// one address are serialised (about 10 ns each), so the block takes the maxima of its
// cells first and issues at most one pair.
template <class R, bool PER> __global__ void __launch_bounds__(256) tracer_src_kernel(ArenaRef ar, adj::TracerAdjArgs<R> A,
                                                                                      int rows, int *dmax) {
    __shared__ int sm[2][256 / WAVE];
    const int x = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y0 = 1 + (int)blockIdx.y * rows, y1 = min(A.ny, y0 + rows - 1);
    int dx = 0, dy = 0;
    if (x <= A.nx)
        for (int y = y0; y <= y1; ++y) {
            const int d = adj::tracer_adj_source<R, PER>(GpuCells<R>{(char *)ar.base, A.P}, A, x, y);
            dx = max(dx, d & 0xFFFF);
            dy = max(dy, d >> 16);
        }
    for (int o = WAVE / 2; o > 0; o >>= 1) {          // every lane takes part
        dx = max(dx, __shfl_xor(dx, o));
        dy = max(dy, __shfl_xor(dy, o));
    }
    const int w = (int)threadIdx.x / WAVE;
    if ((threadIdx.x & (WAVE - 1)) == 0) { sm[0][w] = dx; sm[1][w] = dy; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)blockDim.x / WAVE; ++k) { dx = max(dx, sm[0][k]); dy = max(dy, sm[1][k]); }
        // the maxima only grow: a block whose value does not exceed what it reads skips the atomic
        if (dx > __atomic_load_n(dmax, __ATOMIC_RELAXED)) atomicMax(dmax, dx);
        if (dy > __atomic_load_n(dmax + 1, __ATOMIC_RELAXED)) atomicMax(dmax + 1, dy);
    }
}

// second pass (a) on tiles of GT_X x GT_Y corners, 256 threads, four rows per thread: the
// records of the sources within reach of the tile are staged in LDS, one record kind at a
// time, and each corner scans its window there in tracer_gather's order (rows y = Y+ry down
// to Y-ry, columns x = X+rx down to X-rx), so each sum is the same to the bit.  A source in
// a row outside 1..ny, or (not periodic) a column outside 1..nx, is staged as a corner that
// no (X, Y) matches; periodic columns wrap into 1..nx as there.
// The window shrinks to the tile's own reach: a source at staged column xs, row y whose
// first corner is (cx, cy) can only match the corners (X, Y) = (xs + dx, y + dy) with
// dx in {cx - xs, cx - xs + 1} and dy in {cy - y, cy - y + 1}, so the steps (dx, dy) outside
// the span of those offsets over the staged sources add nothing, and the steps kept are
// summed in the same order.  A corner matches when K - p is 0, 1, 0x10000 or 0x10001 for the
// packed K of (X, Y) and p of the source's first corner (fields of 16 bits, differences
// below 2^15: grids up to 16000 cells a side); bits 0 and 16 of K - p then pick which of the
// source's four terms λ·(wx·wy), formed once when the source is staged, the corner adds.
// Reaches of more than GT_RM - 1 cells, and larger grids, fall back to the scan in global
// memory, as does tiled = 0.  Tiles are GT_X x gt_y<R> corners (64 KiB of LDS at most).
constexpr int GT_X = 64, GT_RM = 5, GT_W = GT_X + 2 * GT_RM;
template <class R> constexpr int gt_y = sizeof(R) == 8 ? 8 : 16;
template <class R> struct GatherTile {
    static constexpr int H = gt_y<R> + 2 * GT_RM;
    R w[4][H][GT_W];                               // the terms of corners (cx, cy), (cx+1, cy), (cx, cy+1), (cx+1, cy+1)
    int cor[H][GT_W];
    int span[4][4];                                // per wave: min / max of the x and y offsets
};
constexpr int GT_NONE = (int)0x80000000;          // K - GT_NONE has bit 31 set: never a match

template <class R, bool PER>
__global__ void __launch_bounds__(256) tracer_gather_uv_tiled_kernel(ArenaRef ar, adj::TracerAdjArgs<R> A, const int *dmax,
                                                                     int tiled) {
    constexpr int GT_Y = gt_y<R>;
    static_assert(GT_X == WAVE && GT_Y % 4 == 0, "one wave per tile row, four waves");
    static_assert(sizeof(GatherTile<R>) <= 65536, "LDS of one workgroup");
    const int rx = dmax[0], ry = dmax[1];
    const GpuCells<R> c{(char *)ar.base, A.P};
    const int X0 = -2 + (int)blockIdx.x * GT_X, Y0 = -1 + (int)blockIdx.y * GT_Y;
    const int lx = (int)threadIdx.x % WAVE, wv = (int)threadIdx.x / WAVE;
    const int X = X0 + lx, nx = A.nx, ny = A.ny;
    if (!tiled || rx + 1 > GT_RM || ry + 1 > GT_RM || nx > 16000 || ny > 16000) {     // uniform over the launch
        if (X <= nx + 3)
            for (int k = 0; k < GT_Y / 4; ++k) {
                const int Y = Y0 + wv + 4 * k;
                if (Y <= ny + 3) adj::tracer_adj_gather_uv<R, PER>(c, A, X, Y, rx, ry);
            }
        return;
    }
    __shared__ GatherTile<R> t;
    const R one = R(1), zero = R(0);
    // the staged block of a kind: rows wv, wv+4, ... and columns lx, lx + 64 of each thread,
    // every load issued before the first is used (loads of the next kind issued before the
    // scan of the current one were slower at 4096^2)
    constexpr int NI = (GatherTile<R>::H + 3) / 4;
    int pv[NI][2];
    R lv[NI][2], xv[NI][2], yv[NI][2];
    auto load = [&](int kind) {
        const uint32_t *rec = kind == 0 ? A.rs : kind == 1 ? A.ru : A.rv;
        const int Rx = rx + (kind ? 1 : 0), Ry = ry + (kind ? 1 : 0);
        const int wa = GT_X + 2 * Rx, ha = GT_Y + 2 * Ry;
#pragma unroll
        for (int a = 0; a < NI; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int i = wv + 4 * a, j = lx + WAVE * b, y = Y0 - Ry + i;
                int x = X0 - Rx + j;
                bool ok = i < ha && j < wa && y >= 1 && y <= ny;
                if (PER) x = ((x - 1) % nx + nx) % nx + 1;
                else ok = ok && x >= 1 && x <= nx;
                pv[a][b] = GT_NONE;
                lv[a][b] = xv[a][b] = yv[a][b] = zero;
                if (ok) {
                    pv[a][b] = c.ati(rec[3], x, y);
                    lv[a][b] = c.at(rec[0], x, y);
                    xv[a][b] = c.at(rec[1], x, y);
                    yv[a][b] = c.at(rec[2], x, y);
                }
            }
    };
    for (int kind = 0; kind < 3; ++kind) {
        const int Rx = rx + (kind ? 1 : 0), Ry = ry + (kind ? 1 : 0);
        const int wa = GT_X + 2 * Rx, ha = GT_Y + 2 * Ry;
        if (kind) __syncthreads();                           // the scans of the previous kind are done
        load(kind);
        int lox = 1 << 20, hix = -(1 << 20), loy = 1 << 20, hiy = -(1 << 20);
#pragma unroll
        for (int a = 0; a < NI; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int i = wv + 4 * a, j = lx + WAVE * b;
                if (i >= ha || j >= wa) continue;
                const int p = pv[a][b];                      // a record's corner packs to a value >= 0
                t.cor[i][j] = p;
                if (p == GT_NONE) continue;
                const R lam = lv[a][b], x0 = xv[a][b], y0 = yv[a][b];
                t.w[0][i][j] = lam * ((one - x0) * (one - y0));
                t.w[1][i][j] = lam * (x0 * (one - y0));
                t.w[2][i][j] = lam * ((one - x0) * y0);
                t.w[3][i][j] = lam * (x0 * y0);
                const int ox = adj::corner_x(p) - (X0 - Rx + j), oy = adj::corner_y(p) - (Y0 - Ry + i);
                lox = min(lox, ox); hix = max(hix, ox);
                loy = min(loy, oy); hiy = max(hiy, oy);
            }
        for (int o = WAVE / 2; o > 0; o >>= 1) {             // every lane takes part
            lox = min(lox, __shfl_xor(lox, o)); hix = max(hix, __shfl_xor(hix, o));
            loy = min(loy, __shfl_xor(loy, o)); hiy = max(hiy, __shfl_xor(hiy, o));
        }
        if (lx == 0) { t.span[wv][0] = lox; t.span[wv][1] = hix; t.span[wv][2] = loy; t.span[wv][3] = hiy; }
        __syncthreads();
        for (int w = 0; w < 4; ++w) {
            lox = min(lox, t.span[w][0]); hix = max(hix, t.span[w][1]);
            loy = min(loy, t.span[w][2]); hiy = max(hiy, t.span[w][3]);
        }
        const int dx0 = max(-Rx, lox), dx1 = min(Rx, hix + 1), dy0 = max(-Ry, loy), dy1 = min(Ry, hiy + 1);
        for (int k = 0; k < GT_Y / 4; ++k) {
            const int ly = wv + 4 * k, Y = Y0 + ly;
            if (X > nx + 3 || Y > ny + 3) continue;
            if (kind == 0 && !(X >= 0 && X <= nx + 1 && Y >= 1 && Y <= ny)) continue;
            const uint32_t K = (uint32_t)adj::pack_corner(X, Y);
            R acc = zero;
            for (int dy = dy0; dy <= dy1; ++dy) {
                const int i = ly - dy + Ry;
                for (int dx = dx0; dx <= dx1; ++dx) {
                    const int j = lx - dx + Rx;
                    const uint32_t d = K - (uint32_t)t.cor[i][j];    // (X - cx) + 0x10000 (Y - cy) mod 2^32
                    const R term = t.w[(d & 1u) | ((d >> 15) & 2u)][i][j];
                    acc = acc + ((d & 0xFFFEFFFEu) == 0u ? term : zero);   // acc is never -0: + 0 leaves it
                }
            }
            c.at(kind == 0 ? A.lsst_out : kind == 1 ? A.gu : A.gv, X, Y) = acc;
        }
    }
}

// second pass (b), one thread per cell X = -2 .. nx+2, Y = -1 .. ny+2
template <class R, bool PER> __global__ void tracer_gather_kernel(ArenaRef ar, adj::TracerAdjArgs<R> A) {
    const int X = -2 + (int)(blockIdx.x * blockDim.x + threadIdx.x), Y = -1 + (int)blockIdx.y;
    if (X > A.nx + 2) return;
    adj::tracer_adj_gather<R, PER>(GpuCells<R>{(char *)ar.base, A.P}, A, X, Y);
}

template <class R> __global__ void snapshot_adj_kernel(R *lu, R *lum, R *lv, R *lvm, size_t n) {
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
        adj::snapshot_adjoint_elem<R>(lu, lum, (long)k);
        adj::snapshot_adjoint_elem<R>(lv, lvm, (long)k);
    }
}

template <class R> __global__ void fold_out_kernel(ArenaRef ar, adj::FoldOutArgs A, R oma, int P) {
    const int x = -2 + (int)(blockIdx.x * blockDim.x + threadIdx.x), y = -1 + (int)blockIdx.y;
    if (x > A.nx + 2) return;
    adj::fold_out_cell<R>(GpuCells<R>{(char *)ar.base, P}, A, oma, x, y);
}

// d = ax·x + ay·y + az·z over n elements (the SSPRK2 stage bookkeeping); y, z nullable
template <class R>
__global__ void lincomb_kernel(R *d, const R *x, R ax, const R *y, R ay, const R *z, R az, size_t n) {
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        adj::lincomb_elem<R>(d, x, ax, y, ay, z, az, (long)k);
}

// the fields of a state snapshot copied in one launch: field f is n16 16-byte words from
// src[f] to dst[f] (16-byte aligned; the host checks)
struct CopySet {
    const uint4 *src[8];
    uint4 *dst[8];
    int n;
    size_t n16;
};
static __global__ void __launch_bounds__(256) copy_fields_kernel(CopySet S) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (int f = 0; f < S.n; ++f)
        for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < S.n16; k += stride) S.dst[f][k] = S.src[f][k];
}

// zero n elements of size 4 (a fill kernel: usable where a memset node is not)
static __global__ void zero_words_kernel(uint32_t *p, size_t n) {
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) p[k] = 0u;
}

}  // namespace swm
