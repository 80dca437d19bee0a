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
// cells first and issues one pair: about two thousand blocks whatever the grid.
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
        atomicMax(dmax, dx);
        atomicMax(dmax + 1, dy);
    }
}

// second pass (a), one thread per corner X = -2 .. nx+3, Y = -1 .. ny+3
template <class R, bool PER> __global__ void tracer_gather_uv_kernel(ArenaRef ar, adj::TracerAdjArgs<R> A, const int *dmax) {
    const int X = -2 + (int)(blockIdx.x * blockDim.x + threadIdx.x), Y = -1 + (int)blockIdx.y;
    if (X > A.nx + 3) return;
    adj::tracer_adj_gather_uv<R, PER>(GpuCells<R>{(char *)ar.base, A.P}, A, X, Y, dmax[0], dmax[1]);
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

// zero n elements of size 4 (a fill kernel: usable where a memset node is not)
static __global__ void zero_words_kernel(uint32_t *p, size_t n) {
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) p[k] = 0u;
}

}  // namespace swm
